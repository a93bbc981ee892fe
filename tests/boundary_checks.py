"""Domain-face ghost fills shared by the CPU-library and GPU tests: the
device's hipPPMExternalBoundary against the oracle's restatement of
ExternalBoundary::SetExternalBoundary, every field bitwise."""
import numpy as np

import boundary_oracle as bo
from enzo_ppm import device, problems
from helpers import field_report

R, O, I, P, S = bo.REFLECTING, bo.OUTFLOW, bo.INFLOW, bo.PERIODIC, bo.SHEARING

EXT_CASES = {
    # name: (active cells, rank, types[6], grid offset, extra domain cells per axis)
    "reflecting": ([6, 5, 7], 3, [R, R, R, R, R, R], [0, 0, 0], [0, 0, 0]),
    "outflow": ([6, 5, 7], 3, [O, O, O, O, O, O], [0, 0, 0], [0, 0, 0]),
    "mixed_inflow": ([6, 5, 7], 3, [I, O, R, I, O, R], [0, 0, 0], [0, 0, 0]),
    "shearing_periodic": ([6, 5, 7], 3, [S, S, P, P, R, O], [0, 0, 0], [0, 0, 0]),
    "inside_domain": ([6, 5, 7], 3, [O, R, R, O, O, O], [4, 0, 0], [4, 0, 3]),
    "one_active_cell": ([1, 5, 4], 3, [R, O, O, R, R, R], [0, 0, 0], [0, 0, 0]),
    "rank2": ([9, 6], 2, [R, O, O, R, P, P], [0, 0, 0], [0, 0, 0]),
}


def held_fields(g):
    """Fields the device holds (Vel3 only in 3-D, gas energy with dual energy)."""
    out = [g.DensNum, g.TENum, g.Vel1Num]
    if g.rank > 1:
        out.append(g.Vel2Num)
    if g.rank > 2:
        out.append(g.Vel3Num)
    if g.dual_energy:
        out.append(g.GENum)
    return out + list(g.colnum)


def run_external_case(lib, name):
    act, rank, types, off, extra = EXT_CASES[name]
    g, glob, _ = problems.random_smooth(act, seed=len(name), ncolours=2, rank=rank)
    rng = np.random.default_rng(1)          # junk in the ghosts
    sl = g.active_slice()
    for f in g.fields:
        keep = f[sl].copy()
        f[...] = rng.normal(size=f.shape)
        f[sl] = keep
    bd = [g.dims[d] + extra[d] + off[d] if d < rank else 1 for d in range(3)]
    inflow = [[100.0 * f + k + 0.5 for k in range(6)] for f in range(g.nfields)]
    b = g.copy()
    bo.external_boundary(b, types, off, bd, inflow=inflow, fields=held_fields(g))
    a = g.copy()
    dg = device.DeviceGrid(a, glob, lib=lib)
    try:
        dg.external_boundary(types, off, bd, inflow=inflow)
        assert dg.download() == 1
    finally:
        dg.close()
    rep = field_report(a, b)
    assert all(r[1] for r in rep), rep
    return sum(int(np.count_nonzero(x != y)) for x, y in zip(a.fields, g.fields))


def run_level_external(lib, types, parts=(2, 1, 2)):
    """hipPPMLevelExternalBoundary on a 12^3 domain cut into siblings (junk
    in every ghost cell) against the oracle's fill of each grid at its
    offset in the domain, every field bitwise."""
    from enzo_ppm.level import Level
    whole, glob, _ = problems.turbulent(12, seed=8, ncolours=2)
    pieces, starts = bo.split_domain(whole, parts)
    rng = np.random.default_rng(5)
    for h in pieces:
        sl = h.active_slice()
        for f in h.fields:
            keep = f[sl].copy()
            f[...] = rng.normal(size=f.shape)
            f[sl] = keep
    gh = whole.ghost
    offs = [[s[d] + gh for d in range(3)] for s in starts]
    bd = [whole.active[d] + 2 * gh for d in range(3)]
    inflow = [[10.0 * f + k for k in range(6)] for f in range(whole.nfields)]
    ref = [h.copy() for h in pieces]
    for h, o in zip(ref, offs):
        bo.external_boundary(h, types, o, bd, inflow=inflow)
    lv = Level([h.copy() for h in pieces], glob, lib=lib)
    try:
        lv.external_boundary(types, offs, bd, inflow=inflow)
        lv.download()
    finally:
        lv.close()
    for a, b in zip(lv.grids, ref):
        rep = field_report(a, b)
        assert all(r[1] for r in rep), rep
