"""Generate the reference data tables used as fixtures by the oracle and the tests.

Reads, as text, two data tables that the reference ships and writes them as plain
whitespace-separated number files:

* ``line_quad_rules.txt`` -- the 60 one-dimensional Gauss rules (orders 1..119)
  stored base64/BSON-encoded in
  ``src/L2-foundation/BoSSS.Foundation/LineQuadRules_bin.txt`` (decoded as in
  ``QuadRuleResource.DecodeFromBase64``, ``QuadRuleResource.cs``).  Format::

      NRULES
      ORDER K
      node weight        (K lines)
      ...

* ``cube_polynomials.txt`` / ``square_polynomials.txt`` -- the orthonormal
  polynomial coefficients printed in ``Cube_polynomials.cs`` (``DefinePolynomials``)
  and ``Square.cs`` (``#region POLY_DEF``), up to total degree ``MAXDEG``.  Format::

      NPOLY D
      NTERMS             (per polynomial)
      coeff e_0 .. e_{D-1}   (NTERMS lines; e_d is the power of x_d)

Run from any directory:  python tests/golden/make_reference_tables.py [REFERENCE_ROOT]
"""
import base64
import os
import re
import struct
import sys

MAXDEG = 6
HERE = os.path.dirname(os.path.abspath(__file__))


def _cstr(b, i):
    j = b.index(0, i)
    return b[i:j].decode(), j + 1


def _bson_doc(b, i):
    n = struct.unpack_from('<i', b, i)[0]
    end = i + n
    i += 4
    out = {}
    while i < end - 1:
        t = b[i]
        i += 1
        k, i = _cstr(b, i)
        if t == 0x01:
            v = struct.unpack_from('<d', b, i)[0]
            i += 8
        elif t == 0x10:
            v = struct.unpack_from('<i', b, i)[0]
            i += 4
        elif t in (0x03, 0x04):
            v, i = _bson_doc(b, i)
        else:
            raise ValueError('unexpected BSON type %d' % t)
        out[k] = v
    return out, end


def decode_line_rules(path):
    txt = open(path, 'rb').read().decode('utf-8-sig').strip()
    doc, _ = _bson_doc(base64.b64decode(txt), 0)
    rules = []
    for key in sorted(doc['Quadrules'], key=int):
        r = doc['Quadrules'][key]
        K = len(r['Item3'])
        nodes = [r['Item2'][str(k)]['0'] for k in range(K)]
        weights = [r['Item3'][str(k)] for k in range(K)]
        rules.append((r['Item1'], nodes, weights))
    return rules


_POLY_RE = re.compile(r'OrthonormalPolynomials\[(\d+)\]\s*=\s*p;')
_COEF_RE = re.compile(r'p\.AddCoeff\(\s*([-+0-9.eE]+)\s*,\s*new int\[\]\s*\{([^}]*)\}\s*\)')


def parse_polynomials(path, D):
    polys = {}
    cur = None
    for line in open(path, encoding='utf-8-sig', errors='replace'):
        m = _POLY_RE.search(line)
        if m:
            cur = int(m.group(1))
            polys[cur] = []
            continue
        m = _COEF_RE.search(line)
        if m and cur is not None:
            exps = [int(x) for x in m.group(2).split(',')]
            assert len(exps) == D
            polys[cur].append((float(m.group(1)), exps))
    out = []
    for i in range(len(polys)):
        terms = polys[i]
        deg = max(sum(e) for _, e in terms)
        if deg > MAXDEG:
            break
        out.append(terms)
    return out


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
    fnd = os.path.join(root, 'src', 'L2-foundation', 'BoSSS.Foundation')
    rules = decode_line_rules(os.path.join(fnd, 'LineQuadRules_bin.txt'))
    with open(os.path.join(HERE, 'line_quad_rules.txt'), 'w') as f:
        f.write('%d\n' % len(rules))
        for order, nodes, weights in rules:
            f.write('%d %d\n' % (order, len(nodes)))
            for x, w in zip(nodes, weights):
                f.write('%r %r\n' % (x, w))
    for name, fn, D in (('cube', 'Cube_polynomials.cs', 3), ('square', 'Square.cs', 2)):
        polys = parse_polynomials(os.path.join(fnd, fn), D)
        with open(os.path.join(HERE, '%s_polynomials.txt' % name), 'w') as f:
            f.write('%d %d\n' % (len(polys), D))
            for terms in polys:
                f.write('%d\n' % len(terms))
                for c, e in terms:
                    f.write('%r %s\n' % (c, ' '.join(str(x) for x in e)))
        print(name, len(polys), 'polynomials up to degree', MAXDEG)
    print(len(rules), 'line rules')


if __name__ == '__main__':
    main()
