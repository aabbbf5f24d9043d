"""CPU audit of the memory-variable addresses of the 3D kernels that index the
slab arrays with plain pointers: the fused step (k_fused3d in
cpml_fused3d.hip) and the time-blocked step (k_tb3d in cpml_tb3d.hip).

Field arrays go through buffer resources whose range check drops
out-of-range accesses; the C-PML memory variables (stored only in the slabs)
do not.  A fault of the fused kernel on its first GPU run came from ring
points at x = -1 or y = -1 whose "inside the grid" test lacked the lower
bound, so Slab::s(-1) = -1 gave negative offsets.  This test restates, for
every launch class, workgroup, lane, wavefront and plane of several grids and
tilings, the predicates under which each kernel reads or writes a memory
variable and the offset it uses, and checks 0 <= offset < the array's size:
a predicate or offset that regresses shows here, on the CPU, before any GPU
run.  Bar: no violation.
"""
import numpy as np
import pytest

import layered_media as lm
from case_inputs import inputs


class Slab:
    def __init__(self, wl, wr, n):
        self.wl, self.wr, self.n, self.w = wl, wr, n, wl + wr

    def inn(self, i):
        i = np.asarray(i)
        return (i >= 0) & (i < self.n) & ((i < self.wl) | (i >= self.n - self.wr))

    def s(self, i):
        i = np.asarray(i)
        return np.where(i < self.wl, i, i - (self.n - self.wr) + self.wl)


def slabs_of(prof, n):
    a, K, ah, Kh = prof[0], prof[2], prof[3], prof[5]
    nz = (a != 0) | (ah != 0) | (K != 1) | (Kh != 1)
    half = (n + 1) // 2
    idx = np.nonzero(nz)[0]
    wl = max([i + 1 for i in idx if i < half], default=0)
    wr = max([n - i for i in idx if i >= half], default=0)
    return Slab(wl, wr, n)


def grid_of(case, prec="f32"):
    if case == "3d_layered":
        d = lm.layered3d_inputs(prec, 4, kmax=3.0)
        nz, ny, nx = d["lam"].shape
    else:
        d = inputs(case, prec, 4)
        nz, ny, nx = d["lam"].shape
    sx, sy, sz = (slabs_of(d[f"prof_{a}"], n) for a, n in (("x", nx), ("y", ny), ("z", nz)))
    mld = (nx + 63) // 64 * 64  # round_up(NX, 256 bytes / 4) in float
    return nx, ny, nz, sx, sy, sz, mld


class Audit:
    def __init__(self, nx, ny, nz, sx, sy, sz, mld):
        self.size = {"x": sx.w * ny * nz, "y": mld * sy.w * nz, "z": mld * ny * sz.w}
        self.bad = 0
        self.checked = 0

    def check(self, axis, mask, off):
        off = np.asarray(off, np.int64)
        sel = np.broadcast_to(mask, off.shape) if np.ndim(mask) else mask
        o = off[sel] if np.ndim(off) else (off if mask else np.array([], np.int64))
        o = np.atleast_1d(o)
        self.checked += o.size
        self.bad += int(((o < 0) | (o >= self.size[axis])).sum())


def clean_tiles_fused(n, B, wl, wr):
    nt = (n + B - 1) // B
    t0 = t1 = 0
    seen = False
    for t in range(nt):
        lo, hi = max(t * B - 1, 0), min(t * B + B, n - 1)
        clean = lo >= wl and hi < n - wr
        if clean and not seen:
            t0, seen = t, True
        if clean:
            t1 = t + 1
        elif seen:
            break
    return t0, t1


def audit_fused(nx, ny, nz, sx, sy, sz, mld, tile_y, kchunk, loose=False):
    """k_fused3d: FO = 62 owned columns, TY = 6 or 14 owned rows + 2 ring rows.
    loose: the first version's "inside" test (no lower bound on i, j)."""
    A = Audit(nx, ny, nz, sx, sy, sz, mld)
    TY = 14 if tile_y > 8 else 6
    FO, NW = 62, TY + 2
    kc = kchunk
    z0 = sz.wl + 1 if sz.wl > 0 else 0
    z1 = nz - sz.wr - 1 if sz.wr > 0 else nz
    if z1 <= z0:
        z0 = z1 = 0
    zn = [(z0 + kc - 1) // kc, (z1 - z0 + kc - 1) // kc, (nz - z1 + kc - 1) // kc]
    b0, b1 = clean_tiles_fused(nx, FO, sx.wl, sx.wr)
    b2, b3 = clean_tiles_fused(ny, TY, sy.wl, sy.wr)
    if b1 == b0 or b3 == b2:
        b0 = b1 = b2 = b3 = 0
    ntx, nty = (nx + FO - 1) // FO, (ny + TY - 1) // TY
    ntz = sum(zn)
    lane = np.arange(64)[None, :]
    wv = np.arange(NW)[:, None]
    for XY in (False, True):
        for Z in (False, True):
            bx, by = b1 - b0, b3 - b2
            per = ntx * nty - bx * by if XY else bx * by
            nc = ntz - zn[1] if Z else zn[1]
            for t in range(per * nc):
                c, r = divmod(t, per)
                tz = (c if c < zn[0] else zn[1] + c) if Z else zn[0] + c
                if not XY:
                    tx, ty = b0 + r % bx, b2 + r // bx
                else:
                    n1, wdt = b2 * ntx, ntx - bx
                    n2 = by * wdt
                    if r < n1:
                        ty, tx = divmod(r, ntx)
                    elif r < n1 + n2:
                        r -= n1
                        ty = b2 + r // wdt
                        cc = r % wdt
                        tx = cc if cc < b0 else cc - b0 + b1
                    else:
                        r -= n1 + n2
                        ty, tx = b3 + r // ntx, r % ntx
                if tz < zn[0]:
                    kb = tz * kc
                    ke = min(kb + kc, z0)
                elif tz < zn[0] + zn[1]:
                    kb = z0 + (tz - zn[0]) * kc
                    ke = min(kb + kc, z1)
                else:
                    kb = z1 + (tz - zn[0] - zn[1]) * kc
                    ke = min(kb + kc, nz)
                i = tx * FO - 1 + lane
                j = ty * TY - 1 + wv
                own = (lane >= 1) & (lane <= FO) & (i < nx) & (wv >= 1) & (wv <= TY) & (j < ny)
                in_xy = (i < nx) & (j < ny)
                if not loose:
                    in_xy = in_xy & (i >= 0) & (j >= 0)
                inx = XY & sx.inn(i)
                iny = XY & sy.inn(j)
                sxs, sys_ = sx.s(i), sy.s(j)
                for s in range(kb - 1, ke + 1):
                    kv = s - 1
                    inside = (0 <= s < nz) & in_xy
                    st = own & (kb <= s < ke)
                    dov = own & (kb <= kv < ke)
                    ox = sxs + sx.w * (j + ny * s)
                    oy = i + mld * (sys_ + sy.w * s)
                    A.check("x", inside & inx, ox)
                    A.check("y", inside & iny, oy)
                    A.check("x", dov & inx, ox - sx.w * ny)
                    A.check("y", dov & iny, oy - mld * sy.w)
                    if Z and 0 <= s < nz and sz.inn(s):
                        A.check("z", inside, i + mld * (j + ny * sz.s(s)))
                    if Z and 0 <= kv < nz and sz.inn(kv):
                        A.check("z", dov, i + mld * (j + ny * sz.s(kv)))
    return A


def clean_tiles_tb(n, B, g, wl, wr):
    nt = (n + B - 1) // B
    t0 = t1 = 0
    seen = False
    for t in range(nt):
        lo, hi = t * B - g, t * B + B - 1 + g
        clean = lo >= wl and hi < n - wr
        if clean and not seen:
            t0, seen = t, True
        if clean:
            t1 = t + 1
        elif seen:
            break
    return t0, t1


def audit_tb3d(nx, ny, nz, sx, sy, sz, mld, NT, NW, ow, kc):
    """k_tb3d: a workgroup of NW wavefronts owns ow columns x NW - 2NT rows."""
    A = Audit(nx, ny, nz, sx, sy, sz, mld)
    ty = NW - 2 * NT
    g = 2 * NT - 1
    z0 = sz.wl + g if sz.wl > 0 else 0
    z1 = nz - sz.wr - g if sz.wr > 0 else nz
    if z1 <= z0:
        z0 = z1 = 0
    zn = [(z0 + kc - 1) // kc, (z1 - z0 + kc - 1) // kc, (nz - z1 + kc - 1) // kc]
    box = list(clean_tiles_tb(nx, ow, NT, sx.wl, sx.wr)) + list(
        clean_tiles_tb(ny, ty, NT, sy.wl, sy.wr))
    if box[1] < box[0]:
        box[0] = box[1] = 0
    if box[3] < box[2]:
        box[2] = box[3] = 0
    n = [(nx + ow - 1) // ow, (ny + ty - 1) // ty, sum(zn)]
    lane = np.arange(64)[None, :]
    wv = np.arange(NW)[:, None]

    def axis_tiles(S, nn, c0, c1):
        return [u for u in range(nn) if (u < c0 or u >= c1)] if S else list(range(c0, c1))

    for SX in (False, True):
        for SY in (False, True):
            for SZ in (False, True):
                for tz in axis_tiles(SZ, n[2], zn[0], zn[0] + zn[1]):
                    if tz < zn[0]:
                        kb = tz * kc
                        ke = min(kb + kc, z0)
                    elif tz < zn[0] + zn[1]:
                        kb = z0 + (tz - zn[0]) * kc
                        ke = min(kb + kc, z1)
                    else:
                        kb = z1 + (tz - zn[0] - zn[1]) * kc
                        ke = min(kb + kc, nz)
                    for tyy in axis_tiles(SY, n[1], box[2], box[3]):
                        for tx in axis_tiles(SX, n[0], box[0], box[1]):
                            i = tx * ow - NT + lane
                            j = tyy * ty - NT + wv
                            col_in, row_in = (i >= 0) & (i < nx), (j >= 0) & (j < ny)
                            in_xy = col_in & row_in
                            own = ((lane >= NT) & (lane < NT + ow) & (i < nx) & (wv >= NT)
                                   & (wv < NT + ty) & (j < ny))
                            inx = SX & sx.inn(i)
                            iny = SY & sy.inn(j)
                            sxi = np.where(inx, sx.s(np.clip(i, 0, nx - 1)), 0)
                            syj = np.where(iny, sy.s(np.clip(j, 0, ny - 1)), 0)

                            def OX(k):
                                return sxi + sx.w * (j + ny * k)

                            def OY(k):
                                return i + mld * (syj + sy.w * k)

                            def OZ(k):
                                return i + mld * (j + ny * sz.s(k))

                            s0, s1 = kb - (2 * NT - 1), ke - 1 + (2 * NT - 1)
                            for s in range(s0, s1 + 1):  # loads of iteration s
                                ks, kv = 0 <= s < nz, 0 <= s - 1 < nz
                                if SX:
                                    if ks:
                                        A.check("x", inx & row_in, OX(s))
                                    if kv:
                                        A.check("x", inx & row_in, OX(s - 1))
                                if SY:
                                    if ks:
                                        A.check("y", iny & col_in, OY(s))
                                    if kv:
                                        A.check("y", iny & col_in, OY(s - 1))
                                if SZ:
                                    if ks and sz.inn(s):
                                        A.check("z", in_xy, OZ(s))
                                    if kv and sz.inn(s - 1):
                                        A.check("z", in_xy, OZ(s - 1))
                                pk, qk = s - 2 * (NT - 1), s - 2 * NT + 1  # level-NT stores
                                for k in (pk, qk):
                                    if kb <= k < ke:
                                        if SX:
                                            A.check("x", own & inx, OX(k))
                                        if SY:
                                            A.check("y", own & iny, OY(k))
                                        if SZ and sz.inn(k):
                                            A.check("z", own, OZ(k))
    return A


GRIDS = ["3d_small", "3d_wide", "3d_layered"]


@pytest.mark.parametrize("tile_y,kchunk", [(4, 8), (16, 48), (16, 5), (8, 3)])
@pytest.mark.parametrize("case", GRIDS)
def test_fused3d_memory_variable_addresses_in_bounds(case, tile_y, kchunk):
    A = audit_fused(*grid_of(case), tile_y, kchunk)
    assert A.checked > 0
    assert A.bad == 0, f"{A.bad} of {A.checked} accesses outside their array"


@pytest.mark.parametrize("NT,NW,ow,kc", [(2, 16, 56, 64), (2, 16, 20, 5), (1, 16, 62, 9),
                                         (2, 8, 33, 16), (1, 8, 8, 200)])
@pytest.mark.parametrize("case", GRIDS)
def test_tb3d_memory_variable_addresses_in_bounds(case, NT, NW, ow, kc):
    A = audit_tb3d(*grid_of(case), NT, NW, ow, kc)
    assert A.checked > 0
    assert A.bad == 0, f"{A.bad} of {A.checked} accesses outside their array"


def test_audit_catches_the_old_fused_predicate():
    """The audit itself: with the old ring predicate (no lower bound on i, j)
    the fused kernel's x-slab offsets go negative, and the audit says so."""
    nx, ny, nz, sx, sy, sz, mld = grid_of("3d_small")

    class Loose(Slab):
        def inn(self, i):
            i = np.asarray(i)
            return (i < self.n) & ((i < self.wl) | (i >= self.n - self.wr))

    A = audit_fused(nx, ny, nz, Loose(sx.wl, sx.wr, sx.n), Loose(sy.wl, sy.wr, sy.n), sz, mld,
                    4, 8, loose=True)
    assert A.bad > 0
