"""Segment bounds of the grouped staged forward projector (k_fp_grouped), on the CPU.

A workgroup serves 64 bins of up to four lines adjacent in slope; per block of FPG_ROWS rows it
stages the columns [sb, sb + 64*CPL) of the frame image, where sb comes from the lines' extreme
rays at the block's first and last rows and CPL from the geometry's group table.  This restates
the table rule (`fpg_groups` in csrc/sk_geom.hip: lines of one frame and sign of cos in slope
order, the spread of x = i*slope + (rho0 + r*drho)/cost over the group's bins and any block,
bounded at the image's first and last block positions) and the kernel's float32 arithmetic
(`fpg_item`: the row hull from row_clip, sb from RN(RN(i*slope) + min intercept); `fps_walk`:
each lane's xbot = RN(RN(i*slope) + icpt) and its pair of columns), and checks that every read
of every lane, row and block lies inside its block's segment.  Results of the kernel itself are
checked bitwise against the other variants on the GPU (test_gpu_batch.py).
"""
import numpy as np
import pytest

f32 = np.float32
ROWS, BINS, LINES, MAXCPL = 16, 64, 4, 4


def frame_lines(angles, N):
    """Steep then flat lines with cost and slope as the geometry builds them (sirtlib.fut:6-7)."""
    out = []
    for a in angles:
        c, s = np.cos(f32(a), dtype=f32), np.sin(f32(a), dtype=f32)
        flat = abs(s) >= abs(c)
        cost, sint = (f32(-s), f32(c)) if flat else (f32(c), f32(-s))
        out.append((flat, cost, f32(-sint / cost)))
    steep = [l for l in out if not l[0]]
    return steep + [l for l in out if l[0]], len(steep)


def group_table(lines, ns, N, R, rho0, drho, rows=ROWS, drift_rows=None):
    """[(lines, CPL, width bound)] per group, or None.  drift_rows: the rows of a block the
    spread is taken over (the rule's rows - 1 drift; other values plant a mistake)."""
    h, nb = N // 2, (R + BINS - 1) // BINS
    drift = rows - 1 if drift_rows is None else drift_rows
    table = []
    for cls in range(4):
        frame, pos = cls >> 1, bool(cls & 1)
        idx = [j for j in (range(ns) if frame == 0 else range(ns, len(lines))) if (lines[j][1] > 0) == pos]
        idx.sort(key=lambda j: float(lines[j][2]))
        for g0 in range(0, len(idx), LINES):
            grp = idx[g0:g0 + LINES]
            spread, inv = 0.0, 0.0
            for b in range(nb):
                r0, r1 = b * BINS, b * BINS + BINS - 1
                for ib in (-h, h - 1):
                    xs = []
                    for j in grp:
                        _, cost, slope = lines[j]
                        inv = max(inv, 1.0 / abs(float(cost)))
                        xs += [i * float(slope) + (rho0 + r * drho) / float(cost)
                               for i in (ib, ib + drift) for r in (r0, r1)]
                    spread = max(spread, max(xs) - min(xs))
            rnd = 4.0 * (abs(rho0) + R * drho) * inv * 2.0**-23 + (h + 16) * 2.0**-22
            w = int(np.ceil(spread + rnd + 0.01)) + 5
            cpl = max(2, (w + 63) // 64)
            if cpl > MAXCPL:
                return None
            table.append((grp, cpl, w))
    return table


def row_clip(slope, icpt, h):
    lo, hi = f32(-h - 2.0), f32(h + 2.0)
    if abs(slope) < 1e-6:
        return -h, h
    ta, tb = f32((lo - icpt) / slope), f32((hi - icpt) / slope)
    tmin = max(min(ta, tb), f32(-h - 4.0))
    tmax = min(max(ta, tb), f32(h + 4.0))
    return max(-h, int(np.floor(tmin)) - 2), min(h, int(np.ceil(tmax)) + 3)


def reads(angles, N, R, rho0, drho, rows=ROWS):
    """Per group of the table: (CPL, width bound, widest segment use, lowest read - sb) over
    every block, lane and row, with the kernel's float32 arithmetic."""
    lines, ns = frame_lines(angles, N)
    rho0, drho = f32(rho0), f32(drho)
    table = group_table(lines, ns, N, R, float(rho0), float(drho), rows)
    assert table is not None
    h, nb = N // 2, (R + BINS - 1) // BINS
    out = []
    for grp, cpl, w in table:
        used, under = 0, 10**9
        for blk in range(nb):
            first, last = blk * BINS, min(blk * BINS + BINS - 1, R - 1)
            r = np.arange(first, first + BINS, dtype=f32)
            icpts = [((rho0 + r * drho) / lines[j][1]).astype(f32) for j in grp]
            hull, sl, ilo = [], [], []
            for j, icp in zip(grp, icpts):
                slope = lines[j][2]
                hull += [row_clip(slope, icp[0], h),
                         row_clip(slope, f32((rho0 + f32(last) * drho) / lines[j][1]), h)]
                sl.append(slope)
                ilo.append(min(icp[0], icp[BINS - 1]))
            i0, i1 = min(a for a, _ in hull), max(b for _, b in hull)
            for ib in range(i0, i1, rows):
                sb = min(int(np.floor(f32(f32(f32(ib + ro) * s) + il))) + h - 1
                         for s, il in zip(sl, ilo) for ro in (0, rows - 1))
                rws = np.arange(ib, ib + rows).astype(f32)
                for j, icp in zip(grp, icpts):
                    tab = (rws * lines[j][2]).astype(f32)
                    col = np.floor((tab[:, None] + icp[None, :]).astype(f32)).astype(np.int64) + h
                    # a lane reads columns floor(xbot) + h - 1 .. + 1 (its pair for either sign)
                    used = max(used, int(col.max()) + 1 - sb + 1)
                    under = min(under, int(col.min()) - 1 - sb)
        out.append((cpl, w, used, under))
    return out


def _uniform(A):
    return (np.arange(A, dtype=np.float64) * np.pi / A).astype(np.float32)


CASES = [
    ("64/90", _uniform(90), 64, 91, -45.0, 1.0),
    ("256/180", _uniform(180), 256, 363, -181.0, 1.0),
    ("off-centre", _uniform(37), 102, 300, -180.5, 0.9),
    ("wide bins", _uniform(23), 40, 30, -29.0, 2.0),
    ("fine bins", _uniform(19), 24, 70, -17.25, 0.5),
    ("shuffled", np.random.default_rng(3).uniform(0, np.pi, 100).astype(np.float32), 128, 181, -90.0, 1.0),
]


@pytest.mark.parametrize("name,angles,N,R,rho0,drho", CASES, ids=[c[0] for c in CASES])
def test_every_read_inside_its_segment(name, angles, N, R, rho0, drho):
    for cpl, w, used, under in reads(angles, N, R, rho0, drho):
        assert under >= 0                 # no read left of the segment start
        assert used <= w <= 64 * cpl      # the width bound covers every read


def test_a_group_table_exists_for_the_benchmark_shape():
    """The 2048^2 / 2048-angle / 2897-bin geometry gets a table, every group at two chunks."""
    lines, ns = frame_lines(_uniform(2048), 2048)
    table = group_table(lines, ns, 2048, 2897, -1448.0, 1.0)
    assert table is not None and len(table) == 513 and {c for _, c, _ in table} == {2}


def test_the_bound_needs_the_block_drift():
    """Leaving out the block's drift over its rows (the spread taken at one row per block
    position) undercounts the width some group's reads need."""
    angles, N, R, rho0, drho = _uniform(90), 64, 91, -45.0, 1.0
    lines, ns = frame_lines(angles, N)
    wrong = group_table(lines, ns, N, R, rho0, drho, drift_rows=0)
    used = [u for _, _, u, _ in reads(angles, N, R, rho0, drho)]
    assert any(w < u for (_, _, w), u in zip(wrong, used))


# ------------------------------------------------------------ exact per-block step (fps_walk)
FP_STEP_BINADES = 16


def tile_binade(x0, x1):
    lo, hi = x0 - 4, x1 + 5
    if lo <= 0 <= hi:
        return -1
    a, b = (lo, hi) if lo > 0 else (-hi, -lo)
    ea, eb = a.bit_length() - 1, b.bit_length() - 1
    return ea if ea == eb else -1


def binade_step(slope, e):
    t = np.ldexp(abs(f32(slope)), 23 - e).astype(f32)
    if t - np.floor(t) == f32(0.5):
        return f32(0.0)
    return np.ldexp(np.rint(t), e - 23).astype(f32)


def exact_blocks(angles, N, R, rho0, drho, rows=ROWS):
    """Every (wave, block) the scalar-unit decision sends down the exact step, with the check
    that it may: each lane's and row's x_bot and x_top lie in the binade and differ by exactly
    D_e, and x_top - floor(x_bot), floor(x_bot) - x_top equal fract(x_bot) + D_e, D_e -
    fract(x_bot) in float32.  Returns (exact blocks, all blocks)."""
    lines, ns = frame_lines(angles, N)
    rho0, drho = f32(rho0), f32(drho)
    table = group_table(lines, ns, N, R, float(rho0), float(drho), rows)
    h, nb = N // 2, (R + BINS - 1) // BINS
    n_exact = n_all = 0
    for grp, _, _ in table:
        for blk in range(nb):
            first, last = blk * BINS, min(blk * BINS + BINS - 1, R - 1)
            r = np.arange(first, first + BINS, dtype=f32)
            hull = []
            icpts = [((rho0 + r * drho) / lines[j][1]).astype(f32) for j in grp]
            for j, icp in zip(grp, icpts):
                hull += [row_clip(lines[j][2], icp[0], h),
                         row_clip(lines[j][2], f32((rho0 + f32(last) * drho) / lines[j][1]), h)]
            i0, i1 = min(a for a, _ in hull), max(b for _, b in hull)
            for j, icp in zip(grp, icpts):
                slope = lines[j][2]
                ic_lo, ic_hi = min(icp[0], icp[63]), max(icp[0], icp[63])
                if not (N <= 8192 and ic_lo > -16384 and ic_hi < 16384):
                    continue
                s_fx = int(np.rint(slope * f32(65536.0)))
                lo_fx, hi_fx = int(np.floor(ic_lo * f32(65536.0))), int(np.ceil(ic_hi * f32(65536.0)))
                for ib in range(i0, i1, rows):
                    n_all += 1
                    a, z = ib * s_fx, (ib + rows - 1) * s_fx
                    e = tile_binade(((min(a, z) + lo_fx) >> 16) - 1, ((max(a, z) + hi_fx) >> 16) + 1)
                    if e < 0 or e >= FP_STEP_BINADES:
                        continue
                    D = binade_step(slope, e)
                    if not D > 0:
                        continue
                    n_exact += 1
                    tab = (np.arange(ib, ib + rows).astype(f32) * slope).astype(f32)
                    xb = (tab[:, None] + icp[None, :]).astype(f32)
                    xt = (xb + slope).astype(f32)
                    for x in (xb, xt):
                        m = np.abs(x)
                        assert np.all(m >= 2.0**e) and np.all(m < 2.0**(e + 1)), (j, ib, e)
                        assert np.all(np.sign(x) == np.sign(xb[0, 0]))
                    assert np.all(np.abs((xt - xb).astype(f32)) == D)
                    fx = np.floor(xb).astype(f32)
                    fr = (xb - fx).astype(f32)
                    if slope >= 0:
                        assert np.array_equal((xt - fx).astype(f32), (fr + D).astype(f32))
                    else:
                        assert np.array_equal((fx - xt).astype(f32), (D - fr).astype(f32))
    return n_exact, n_all


@pytest.mark.parametrize("name,angles,N,R,rho0,drho", CASES[:2] + CASES[3:4],
                         ids=[CASES[0][0], CASES[1][0], CASES[3][0]])
def test_exact_step_blocks_are_exact(name, angles, N, R, rho0, drho):
    n_exact, n_all = exact_blocks(angles, N, R, rho0, drho)
    assert n_all > 0
    if N >= 256:
        assert n_exact > 0  # the step applies somewhere (binades wide enough for a block)
