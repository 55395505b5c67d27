#!/usr/bin/env python3
"""Re-encode the marching-cubes lookup tables into the flat layout used here.

The triangulations themselves must equal the reference's, so this script reads
them (as text) from a checkout of neurolabusc/nii2mesh:

* ``src/oldcubes.c``    ``triTable[256][16]``   (Bourke/Bloyd classic table)
* ``src/LookUpTable.h`` ``cases``/``test*``/``tiling*``/``subconfig13``/
  ``casesClassic``   (Lewiner MC33 tables)

and writes ``mc_tables.h`` with a different, device-friendly layout:

* classic: one 64-bit word per cube configuration, edge numbers packed 4 bits
  each, low nibble first, 0xF terminates (a configuration has <= 5 triangles,
  15 nibbles);
* MC33: every tiling row of every tiling table becomes one *row* of a single
  flat byte array of edge codes (0..11 = cube edges, 12 = cube-centre vertex),
  addressed by ``row -> (offset, triangle count)``.  A family base constant
  gives the first row of each tiling table; rows are laid out
  ``config * nsub + sub`` for the tables with a sub-configuration index.

Usage: ``python gen_mc_tables.py /path/to/nii2mesh/src > mc_tables.h``.
The output is committed; this script is only needed to regenerate it.
"""
import re
import sys


def _strip(text):
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return re.sub(r'//[^\n]*', '', text)


def _parse_tables(text, qualifier):
    text = _strip(text)
    out = {}
    pat = re.compile(qualifier + r'\s+(\w+)((?:\[\d+\])+)\s*=\s*\{')
    for m in pat.finditer(text):
        name = m.group(1)
        dims = [int(d) for d in re.findall(r'\[(\d+)\]', m.group(2))]
        depth, i = 1, m.end()
        while depth:
            depth += {'{': 1, '}': -1}.get(text[i], 0)
            i += 1
        nums = [int(v, 0) for v in re.findall(r'-?0x[0-9a-fA-F]+|-?\d+', text[m.end():i])]
        total = 1
        for d in dims:
            total *= d
        if len(nums) != total:
            raise ValueError('%s: %d values, expected %d' % (name, len(nums), total))
        out[name] = (dims, nums)
    return out


# (family symbol, table name, triangles per row or None = count to -1, uses centre vertex)
FAMILIES = [
    ('T1', 'tiling1', 1), ('T2', 'tiling2', 2),
    ('T3_1', 'tiling3_1', 2), ('T3_2', 'tiling3_2', 4),
    ('T4_1', 'tiling4_1', 2), ('T4_2', 'tiling4_2', 6),
    ('T5', 'tiling5', 3),
    ('T6_1_1', 'tiling6_1_1', 3), ('T6_1_2', 'tiling6_1_2', 9), ('T6_2', 'tiling6_2', 5),
    ('T7_1', 'tiling7_1', 3), ('T7_2', 'tiling7_2', 5), ('T7_3', 'tiling7_3', 9),
    ('T7_4_1', 'tiling7_4_1', 5), ('T7_4_2', 'tiling7_4_2', 9),
    ('T8', 'tiling8', 2), ('T9', 'tiling9', 4),
    ('T10_1_1', 'tiling10_1_1', 4), ('T10_1_1_', 'tiling10_1_1_', 4),
    ('T10_1_2', 'tiling10_1_2', 8), ('T10_2', 'tiling10_2', 8), ('T10_2_', 'tiling10_2_', 8),
    ('T11', 'tiling11', 4),
    ('T12_1_1', 'tiling12_1_1', 4), ('T12_1_1_', 'tiling12_1_1_', 4),
    ('T12_1_2', 'tiling12_1_2', 8), ('T12_2', 'tiling12_2', 8), ('T12_2_', 'tiling12_2_', 8),
    ('T13_1', 'tiling13_1', 4), ('T13_1_', 'tiling13_1_', 4),
    ('T13_2', 'tiling13_2', 6), ('T13_2_', 'tiling13_2_', 6),
    ('T13_3', 'tiling13_3', 10), ('T13_3_', 'tiling13_3_', 10),
    ('T13_4', 'tiling13_4', 12),
    ('T13_5_1', 'tiling13_5_1', 6), ('T13_5_2', 'tiling13_5_2', 10),
    ('T14', 'tiling14', 4),
    ('CLASSIC', 'casesClassic', None),
]
# tilings drawn with a cube-centre vertex (code 12) in the MC33 decision tree
CENTRE = {'T6_1_2', 'T7_3', 'T10_2', 'T10_2_', 'T12_2', 'T12_2_', 'T13_3', 'T13_3_', 'T13_4'}
TESTS = ['test3', 'test4', 'test6', 'test7', 'test10', 'test12', 'test13', 'subconfig13']


def _fmt(name, ctype, values, per_line=24):
    lines = []
    for i in range(0, len(values), per_line):
        lines.append('  ' + ','.join(str(v) for v in values[i:i + per_line]) + ',')
    return 'N2M_TABLE_QUAL %s %s[%d] = {\n%s\n};\n' % (ctype, name, len(values), '\n'.join(lines))


def main(srcdir):
    lut = _parse_tables(open(srcdir + '/LookUpTable.h').read(), r'static\s+const\s+char')
    old = _parse_tables(open(srcdir + '/oldcubes.c').read(), r'int')
    tri_dims, tri = old['triTable']
    assert tri_dims == [256, 16]
    # classic table: packed nibbles
    packed = []
    for c in range(256):
        row = tri[c * 16:(c + 1) * 16]
        n = row.index(-1) if -1 in row else 16
        assert n % 3 == 0 and n <= 15
        word = 0
        for k in range(15):
            word |= (row[k] if k < n else 0xF) << (4 * k)
        packed.append('0x%016xull' % word)
    cls_dims, cls = lut['casesClassic']
    same = cls == tri
    # MC33 rows
    bases, offs, ntris, edges, row_has_c = [], [], [], [], []
    for sym, name, n in FAMILIES:
        dims, vals = lut[name]
        width = dims[-1]
        nrows = len(vals) // width
        bases.append((sym, len(offs), nrows))
        for r in range(nrows):
            row = vals[r * width:(r + 1) * width]
            k = n if n is not None else (row.index(-1) // 3 if -1 in row else width // 3)
            codes = row[:3 * k]
            assert all(0 <= e <= 12 for e in codes), (name, r, codes)
            if sym not in CENTRE:
                assert 12 not in codes, (name, r)
            offs.append(len(edges))
            ntris.append(k)
            edges.extend(codes)
            # add_c_vertex() runs for every row of a centre family (MarchingCubes.c:523-685)
            row_has_c.append(1 if sym in CENTRE else 0)
    case_dims, case_vals = lut['cases']
    case = [case_vals[2 * i] for i in range(256)]
    config = [max(case_vals[2 * i + 1], 0) for i in range(256)]
    out = []
    out.append('/* Generated by gen_mc_tables.py from the neurolabusc/nii2mesh marching-cubes\n'
               ' * tables (src/oldcubes.c triTable, src/LookUpTable.h); do not edit.\n'
               ' * The triangulations are identical to the reference; only the layout differs. */\n')
    out.append('#ifndef N2M_MC_TABLES_H\n#define N2M_MC_TABLES_H\n#include <stdint.h>\n')
    out.append('#ifndef N2M_TABLE_QUAL\n#define N2M_TABLE_QUAL static const\n#endif\n')
    out.append('/* casesClassic (MC33 "-o 1") equals the classic table: %s */\n' % ('yes' if same else 'NO'))
    out.append('#define N2M_CLASSIC_EQUALS_MC33_ORIGINAL %d\n' % (1 if same else 0))
    out.append('/* classic: bit p of the configuration = corner p below the isolevel */\n')
    out.append(_fmt('n2m_classic_tri', 'uint64_t', packed, 4))
    out.append(_fmt('n2m_lw_case', 'uint8_t', case))
    out.append(_fmt('n2m_lw_config', 'uint8_t', config))
    for t in TESTS:
        dims, vals = lut[t]
        out.append(_fmt('n2m_lw_' + t, 'int8_t', vals))
    for sym, base, nrows in bases:
        out.append('#define N2M_LW_%s %d /* %d rows */\n' % (sym, base, nrows))
    out.append('#define N2M_LW_NROWS %d\n#define N2M_LW_NEDGES %d\n' % (len(offs), len(edges)))
    out.append(_fmt('n2m_lw_row_off', 'uint16_t', offs))
    out.append(_fmt('n2m_lw_row_ntri', 'uint8_t', ntris))
    out.append(_fmt('n2m_lw_row_centre', 'uint8_t', row_has_c))
    out.append(_fmt('n2m_lw_edges', 'uint8_t', edges, 30))
    out.append('#endif\n')
    sys.stdout.write('\n'.join(out))


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference/src')
