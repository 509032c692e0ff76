"""Extract the constant tables of the reference (as data) into tests/golden/ref_tables.json.

The reference (mnivan/TOP_GPU_on_matlab) ships no tests, fixtures or golden vectors; the
only data it holds for the hot path are the constant tables written inline in TOP_GPU.m:

  * Solving_Operator4MultiGridRestrictionAndInterpolation  (TOP_GPU.m:2165-2328): RI, span 2 / 4
  * Solving_SubEleNodMat                                     (TOP_GPU.m:1938-2017): span 2 / 4
  * reOrdering                                               (TOP_GPU.m:1051)
  * FEA_VoxelBasedElementStiffnessMatrix                     (TOP_GPU.m:1465-1503): the
    coefficient vector C and the 6x6 blocks A11..B23 of the closed-form Ke

This script reads those numbers out of the MATLAB source text and writes them as JSON
(Ke is evaluated from the blocks for nu = 0.3 and cellSize = 1, using the block layout of
TOP_GPU.m:1499-1500).  It needs /root/reference and is run once, in the build container:

    python tests/golden/make_ref_tables.py
"""
import json
import os
import re
from fractions import Fraction

import numpy as np

REF = os.environ.get('TOPGPU_REFERENCE', '/root/reference/TOP_GPU.m')
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'ref_tables.json')


def _matrix_after(text, start_pat, end_pat='];'):
    i = text.index(start_pat) + len(start_pat)
    j = text.index(end_pat, i)
    body = text[i:j].replace('...', ' ')
    rows = []
    for line in re.split(r'[;\n]', body):
        toks = line.split()
        if toks:
            rows.append(toks)
    return rows


def _numeric(rows):
    return [[float(Fraction(t)) for t in r] for r in rows]


def main():
    text = open(REF).read()
    fn_ri = text.index('function ss = Solving_Operator4MultiGridRestrictionAndInterpolation')
    ri_text = text[fn_ri:]
    ri2 = _numeric(_matrix_after(ri_text[ri_text.index('case 2'):], 'ss = ['))
    ri4 = _numeric(_matrix_after(ri_text[ri_text.index('case 4'):], 'ss = ['))
    fn_sub = text.index('function val = Solving_SubEleNodMat')
    sub_text = text[fn_sub:]
    sub2 = [[int(t) for t in r] for r in _matrix_after(sub_text[sub_text.index('case 2'):], 'val = [')]
    sub4 = [[int(t) for t in r] for r in _matrix_after(sub_text[sub_text.index('case 4'):], 'val = [')]
    m = re.search(r'reOrdering = \[([0-9 ]+)\]', text)
    reord = [int(t) for t in m.group(1).split()]

    fn_ke = text.index('function KE = FEA_VoxelBasedElementStiffnessMatrix')
    ke_text = text[fn_ke:text.index('function FEA_ApplyBoundaryCondition')]
    cvals = [Fraction(t) for t in re.search(r'C = \[([^\]]+)\]', ke_text).group(1).split()]

    def block(name):
        rows = _matrix_after(ke_text, name + ' = [')
        out = np.zeros((6, 6))
        for r, row in enumerate(rows):
            for c, tok in enumerate(row):
                if tok == '0':
                    continue
                sign = -1.0 if tok.startswith('-') else 1.0
                k = int(re.search(r'C\((\d)\)', tok).group(1))
                out[r, c] = sign * float(cvals[k - 1])
        return out

    A = {n: block(n) for n in ('A11', 'A22', 'A12', 'A13', 'A14', 'A23')}
    B = {n: block(n) for n in ('B11', 'B22', 'B12', 'B13', 'B14', 'B23')}

    def layout(M, p):
        return np.block([
            [M[p + '11'], M[p + '12'], M[p + '13'], M[p + '14']],
            [M[p + '12'].T, M[p + '22'], M[p + '23'], M[p + '13'].T],
            [M[p + '13'].T, M[p + '23'].T, M[p + '22'], M[p + '12'].T],
            [M[p + '14'].T, M[p + '13'], M[p + '12'], M[p + '11']]])

    nu = 0.3
    ke = 1 / (1 + nu) / (2 * nu - 1) * (layout(A, 'A') + nu * layout(B, 'B'))

    data = {
        'source': 'TOP_GPU.m constant tables (mnivan/TOP_GPU_on_matlab)',
        'RI_span2': ri2, 'RI_span4': ri4,
        'SubEleNodMat_span2': sub2, 'SubEleNodMat_span4': sub4,
        'reOrdering': reord,
        'Ke_nu0.3_cell1': ke.tolist(),
    }
    with open(OUT, 'w') as f:
        json.dump(data, f)
    print('wrote', OUT, {k: np.shape(v) for k, v in data.items() if k != 'source'})


if __name__ == '__main__':
    main()
