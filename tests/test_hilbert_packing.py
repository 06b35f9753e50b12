"""Operand packing of the matrix-core Hilbert-space kernels, on the CPU.  Each kernel's K loop is
emulated in NumPy with the instruction's lane layout -- the one the kernel relies on -- and the
state kept in the accumulator layout; the emulated product must be H_eff rho.  A wrong row
order, a transposed tile or a mis-indexed lane fails these."""
import numpy as np
import pytest

from gqis import hilbert, workloads
from gqis.solver import _drive_map, _prepare_model


def _model(n, seed=3):
    w = workloads.random_n_level(n, seed=seed, n_eps=2, n_amp=2)
    model = _prepare_model(w["H"], w["Drive"], w["Col_Ops"], w["mean_operator"],
                           sweep_symbols=tuple(w["var_arrays"]), const_values=None,
                           runtime_consts=None, keep_symbolic_consts=None,
                           auto_runtime_consts=False, rho0=None, ode_size=None)
    return hilbert.extract(model, _drive_map(model["Drive"]))


def _rbig(rho):
    return np.concatenate([rho.real, rho.imag], axis=0)      # (2N, N): [Re; Im]


def _xbig(m, rho):
    x = m.h_eff @ rho                                       # H_eff rho
    return np.concatenate([x.real, x.imag], axis=0)


@pytest.mark.parametrize("n", [8, 16, 32])
def test_fp64_matrix_core_packing(n):
    # v_mfma_f64_16x16x4f64: A lane l = A[l % 16][l // 16], B lane l = B[l // 16][l % 16],
    # result register r of lane l = D[4 r + l // 16][l % 16] (measured on gfx950)
    m = _model(n)
    rng = np.random.default_rng(n)
    rho = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    perm = hilbert.dmfma_order(n)
    assert sorted(perm) == list(range(2 * n))
    pack = hilbert.pack_a_dmfma(m)                             # [To][T][r][lane]
    tr, tc, pts = n // 8, (2 if n == 32 else 1), (2 if n == 8 else 1)
    rb = _rbig(rho)
    # the wavefront's 16-column tiles: N = 8 puts two points (here the same rho twice) side
    # by side, N = 32 one point in two tiles
    cols = np.concatenate([rb] * pts, axis=1) if pts == 2 else rb
    # lane (c, g) of column tile ct holds, in row tile T and register r, physical row
    # 16 T + 4 r + g = logical Rbig row perm[...]
    x = np.zeros((tc, tr, 4, 64))
    for ct in range(tc):
        for t in range(tr):
            for r in range(4):
                for lane in range(64):
                    c, g = lane % 16, lane // 16
                    x[ct, t, r, lane] = cols[perm[16 * t + 4 * r + g], 16 * ct + c]
    acc = np.zeros((tr, tc, 4, 64))
    for t in range(tr):
        for r in range(4):
            for to in range(tr):
                a = pack[to, t, r]
                for ct in range(tc):
                    A = np.zeros((16, 4))
                    B = np.zeros((4, 16))
                    for lane in range(64):
                        A[lane % 16, lane // 16] = a[lane]
                        B[lane // 16, lane % 16] = x[ct, t, r, lane]
                    D = A @ B
                    for lane in range(64):
                        for rr in range(4):
                            acc[to, ct, rr, lane] += D[4 * rr + lane // 16, lane % 16]
    xb = _xbig(m, rho)
    want = np.concatenate([xb] * pts, axis=1) if pts == 2 else xb
    for to in range(tr):
        for ct in range(tc):
            for rr in range(4):
                for lane in range(64):
                    c, g = lane % 16, lane // 16
                    got = acc[to, ct, rr, lane]
                    assert abs(got - want[perm[16 * to + 4 * rr + g], 16 * ct + c]) < 1e-12


def test_fp64_order_pairs_real_and_imaginary_parts():
    # each lane group holds the real and imaginary parts of the same N/4 elements
    for n in (8, 16, 32):
        perm = hilbert.dmfma_order(n)
        q = n // 4
        for g in range(4):
            slots = [perm[16 * (m // 4) + 4 * (m % 4) + g] for m in range(n // 2)]
            re, im = slots[:q], slots[q:]
            assert all(r < n for r in re) and [i - n for i in im] == re
            assert re == list(range(g * q, (g + 1) * q))


@pytest.mark.parametrize("n", [8, 16, 32])
def test_fp32_and_split_matrix_core_packing(n):
    # v_mfma_f32_32x32x2f32: A lane l = A[l % 32][l // 32], B lane l = B[l // 32][l % 32],
    # result element e of lane l = D[8 (e // 4) + 4 (l // 32) + e % 4][l % 32]; the K step s
    # takes Rbig row kappa(s, h) = 8 (s // 4) + 4 h + s % 4 from lane half h.  The split form
    # (K blocks of 16: A lane (r, h) element j = A[r][8 h + j], B likewise) must give the same
    # product from its hi + lo parts, to FP16 rounding of the lo part.
    m = _model(n)
    rng = np.random.default_rng(n + 1)
    rho = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    pts, tiles = 32 // n, (2 * n + 31) // 32
    rb = np.concatenate([_rbig(rho)] * pts, axis=1)            # 32 columns
    xb = np.concatenate([_xbig(m, rho)] * pts, axis=1)

    def kappa(s, h):
        return 8 * (s // 4) + 4 * h + s % 4
    state = np.zeros((n, 64))                                   # x[s] of lane (c, h)
    for s in range(n):
        for lane in range(64):
            state[s, lane] = rb[kappa(s, lane // 32), lane % 32]
    pack = hilbert.pack_a(m)                                    # [tile][step][lane]
    hi, lo = hilbert.pack_a_split(m, "f16")                     # [tile][block][lane][8]
    for t in range(tiles):
        acc = np.zeros((32, 32))
        acc_split = np.zeros((32, 32))
        for s in range(n):
            A = np.zeros((32, 2))
            B = np.zeros((2, 32))
            for lane in range(64):
                A[lane % 32, lane // 32] = pack[t, s, lane]
                B[lane // 32, lane % 32] = state[s, lane]
            acc += A @ B
        for b in range(n // 8):
            A = np.zeros((32, 16))
            B = np.zeros((16, 32))
            for lane in range(64):
                for j in range(8):
                    A[lane % 32, 8 * (lane // 32) + j] = hi[t, b, lane, j] + lo[t, b, lane, j]
                    B[8 * (lane // 32) + j, lane % 32] = state[8 * b + j, lane]
            acc_split += A @ B
        for lane in range(64):
            for e in range(16):
                row = 32 * t + 8 * (e // 4) + 4 * (lane // 32) + e % 4
                if row < 2 * n:
                    want = xb[row, lane % 32]
                    assert abs(acc[8 * (e // 4) + 4 * (lane // 32) + e % 4, lane % 32] - want) < 1e-12
                    assert abs(acc_split[8 * (e // 4) + 4 * (lane // 32) + e % 4, lane % 32]
                               - want) < 1e-5 * max(1.0, np.max(np.abs(xb)))


def _mfma16(a, b):
    """v_mfma_f32_16x16x32 on lane-held operands a, b [64][8]: A lane l = A[l % 16][8 (l // 16)
    + e], B lane l = B[8 (l // 16) + e][l % 16]; returns D [16][16]."""
    A = np.zeros((16, 32))
    B = np.zeros((32, 16))
    for lane in range(64):
        for e in range(8):
            A[lane % 16, 8 * (lane // 16) + e] = a[lane, e]
            B[8 * (lane // 16) + e, lane % 16] = b[lane, e]
    return A @ B


@pytest.mark.parametrize("n", [16, 32])
@pytest.mark.parametrize("kind", ["f16", "bf16"])
def test_two_sided_split_packing(n, kind):
    # kernels/hilbert_split.hip: lane (c, g) = (l % 16, l // 16) holds element (16 rb + 4 g + r,
    # 16 cb + c) of each real part in register r of tile (P, rb, cb); the state values of K
    # block kb and column block cb are its B operand, and -- for a Hermitian rho -- the A
    # operand of the second product's row block cb.  The second product's coefficients (B
    # operands) are the first's (A operands) for P = 0 and their negation for P = 1, so the
    # kernel loads the first set only and subtracts the P = 1 second products.  Both products
    # together must give -i H_eff rho + i rho H_eff^+ in the state's layout.
    m = _model(n)
    rng = np.random.default_rng(n + 7)
    z = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    rho = (z + z.conj().T) / (2 * n)
    want = -1j * (m.h_eff @ rho) + 1j * (rho @ m.h_eff.conj().T)
    want_p = (want.real, want.imag)
    rho_p = (rho.real, rho.imag)
    nb = n // 16
    lhi, llo, rhi, rlo = hilbert.pack_two_sided(m, kind)
    for kb in range(nb):
        for b in range(nb):
            np.testing.assert_array_equal(rhi[0, kb, b], lhi[0, b, kb])
            np.testing.assert_array_equal(rlo[0, kb, b], llo[0, b, kb])
            np.testing.assert_array_equal(rhi[1, kb, b], -lhi[1, b, kb])
            np.testing.assert_array_equal(rlo[1, kb, b], -llo[1, b, kb])
    half = (np.float16 if kind == "f16" else None)

    def rnd(x):
        return x.astype(half).astype(np.float64) if half else hilbert._round_half(x, kind)
    # S[kb][cb][lane][e]: part e // 4, row 16 kb + 4 g + e % 4, column 16 cb + c
    s = np.zeros((nb, nb, 64, 8))
    for kb in range(nb):
        for cb in range(nb):
            for lane in range(64):
                c, g = lane % 16, lane // 16
                for e in range(8):
                    s[kb, cb, lane, e] = rho_p[e // 4][16 * kb + 4 * g + e % 4, 16 * cb + c]
    sh = rnd(s)
    sl = rnd(s - sh)
    scale = np.max(np.abs(want))
    sign = (1.0, -1.0)
    for p in range(2):
        for rb in range(nb):
            for cb in range(nb):
                exact = np.zeros((16, 16))
                split = np.zeros((16, 16))
                for kb in range(nb):
                    lf = lhi[p, rb, kb] + llo[p, rb, kb]     # first product, row block rb
                    rf = lhi[p, cb, kb] + llo[p, cb, kb]     # second product, column block cb
                    exact += _mfma16(lf, s[kb, cb]) + sign[p] * _mfma16(s[kb, rb], rf)
                    split += (_mfma16(llo[p, rb, kb], sh[kb, cb]) + _mfma16(lhi[p, rb, kb], sl[kb, cb])
                              + _mfma16(lhi[p, rb, kb], sh[kb, cb]))
                    split += sign[p] * (_mfma16(sh[kb, rb], llo[p, cb, kb]) + _mfma16(sl[kb, rb], lhi[p, cb, kb])
                                        + _mfma16(sh[kb, rb], lhi[p, cb, kb]))
                for lane in range(64):
                    c, g = lane % 16, lane // 16
                    for r in range(4):
                        w = want_p[p][16 * rb + 4 * g + r, 16 * cb + c]
                        # D[4 g + r][c] lands in register r of lane (c, g); hi + lo keep 22
                        # (FP16) or 16 (BF16) significant bits
                        tol = (2e-6 if kind == "f16" else 5e-5) * scale
                        assert abs(exact[4 * g + r, c] - w) < tol
                        assert abs(split[4 * g + r, c] - w) < tol
