"""GPU parity of the HIP projector and solver against the CPU oracle (FP32 reference
semantics and FP64), on the reference's anisotropic test phantom and small variants."""
import numpy as np
import pytest

from conftest import announce, rel_l2
from prc_mlem_amd import phantoms

pytestmark = pytest.mark.gpu
SPACING = (1.1, 1.6, 2.3)
ISOTOPES = ["11C", "13N", "15O", "18F", "64Cu", "68Ga", "82Rb", "124I"]


@pytest.fixture(scope="module")
def phantom(oracle_mod):
    rho, x, y, spacing = phantoms.test_phantom()
    rho_p = np.maximum(rho, np.float32(1e-3))
    return rho_p, x, y, rho_p >= np.float32(0.05)


@pytest.fixture(scope="module")
def pair(prc, oracle_mod, phantom):
    """GPU projector and the FP32 oracle built on the GPU's own survival table."""
    rho, x, y, sup = phantom
    announce("build GPU projector 16x14x12 68Ga rays=32")
    A = prc.Projector(rho, SPACING, "68Ga", 32)
    table, _ = A.table
    O = oracle_mod.Projector(rho, SPACING, "68Ga", 32, table=table)
    return A, O


def test_device_range_table_matches_host_table(prc, oracle_mod, pair):
    # Device expf/powf vs host libm; the reference bound vs FP64 is 2e-5 (tests/test_rdp.py:24).
    A, _ = pair
    dev, scale = A.table
    host, hscale = oracle_mod.range_table("68Ga")
    assert scale == pytest.approx(float(hscale), rel=0, abs=0)
    assert np.max(np.abs(dev - host)) <= 2e-7
    r = np.arange(4097, dtype=np.float64) / scale
    exact = np.array([oracle_mod.survival_f64("68Ga", v) for v in r])
    assert np.max(np.abs(dev - exact)) <= 2e-6


def test_normalisation_matches_oracle(pair, phantom):
    A, O = pair
    _, _, _, sup = phantom
    W = A.W
    w_ref = np.zeros(W.size)
    w_ref[O.active] = O.W
    np.testing.assert_array_equal(W[~sup], 0)
    # Same table, same march arithmetic and summation order: agreement to float rounding.
    np.testing.assert_allclose(W.ravel(), w_ref, rtol=2e-6, atol=0)


def test_forward_adjoint_match_oracle(pair, phantom):
    A, O = pair
    _, x, y, sup = phantom
    announce("forward/adjoint 16x14x12")
    ax, aty = A.forward(x), A.adjoint(y)
    assert rel_l2(ax, O.forward(x)) <= 1e-6
    assert rel_l2(aty, O.adjoint(y)) <= 1e-6


def test_identity_conservation_and_support(pair, phantom):
    # tests/test_solver.py:180-190, with a 1e-6 bound.
    A, _ = pair
    _, x, y, sup = phantom
    ax, aty = A.forward(x), A.adjoint(y)
    lhs = np.sum(ax.astype(np.float64) * y)
    rhs = np.sum(x.astype(np.float64) * aty)
    assert abs(lhs / rhs - 1) <= 1e-6
    assert abs(ax.sum(dtype=np.float64) / x.sum(dtype=np.float64) - 1) <= 1e-6
    np.testing.assert_array_equal(ax[~sup], 0)
    at_one = A.adjoint(np.ones_like(x))
    np.testing.assert_allclose(at_one[sup], 1, rtol=3e-6, atol=3e-6)
    np.testing.assert_array_equal(at_one[~sup], 0)


@pytest.mark.parametrize("iso", ISOTOPES)
def test_table_vs_exact_range_all_isotopes(prc, phantom, iso):
    # tests/test_solver.py:201-214: table lookup vs analytic survival per step.
    rho, x, y, sup = phantom
    announce(f"isotope {iso}: table and exact projectors")
    A = prc.Projector(rho, SPACING, iso, 32)
    E = prc.Projector(rho, SPACING, iso, 32, exact_range=True)
    ax, ex = A.forward(x), E.forward(x)
    aty, ety = A.adjoint(y), E.adjoint(y)
    np.testing.assert_allclose(ax, ex, rtol=1e-5, atol=2e-7)
    np.testing.assert_allclose(aty, ety, rtol=1e-5, atol=2e-7)
    lhs = np.sum(ax.astype(np.float64) * y)
    rhs = np.sum(x.astype(np.float64) * aty)
    assert abs(lhs / rhs - 1) <= 1e-6
    assert abs(ax.sum(dtype=np.float64) / x.sum(dtype=np.float64) - 1) <= 1e-6


@pytest.mark.parametrize("beta,accel", [(0.0, False), (0.1, True), (0.2, False), (0.0, True)])
def test_solver_per_iteration_matches_oracle(prc, pair, phantom, beta, accel):
    # Bar: FP32 GPU vs the FP32 oracle within 1e-5 relative L2 per iteration.
    A, O = pair
    _, x, _, sup = phantom
    announce(f"solver beta={beta} accel={accel}")
    _, hist = O.solve(x, 6, beta, 20.0, accel, history=True)
    s = prc.Solver(A, x, beta, 20.0, accel)
    for it, ref in enumerate(hist, 1):
        s.iterate(1)
        est = s.estimate()
        assert rel_l2(est, ref) <= 1e-5, f"iteration {it}"
        np.testing.assert_array_equal(est[~sup], 0)
    tot, support_sum = s.sums()
    assert abs(tot / support_sum - 1) <= 1e-6


def test_mixed_precision_against_fp64(prc, oracle_mod, phantom):
    rho, x, y, sup = phantom
    announce("mixed precision projector")
    M = prc.Projector(rho, SPACING, "68Ga", 32, precision="mixed")
    O64 = oracle_mod.Projector(rho, SPACING, "68Ga", 32, prec=64)
    ref = O64.solve(x, 6, 0.03, 20.0, True)
    out = prc.correct(M, x, 6, 0.03, 20.0, True)
    assert rel_l2(out, ref) <= 1e-3
    ax, aty = M.forward(x), M.adjoint(y)
    lhs = np.sum(ax.astype(np.float64) * y)
    rhs = np.sum(x.astype(np.float64) * aty)
    assert abs(lhs / rhs - 1) <= 1e-6
    assert abs(ax.sum(dtype=np.float64) / x.sum(dtype=np.float64) - 1) <= 1e-6


def test_batch_equals_single_runs(prc, phantom):
    rho, x, y, sup = phantom
    announce("batched correction of three volumes")
    A1 = prc.Projector(rho, SPACING, "68Ga", 32)
    rho2 = np.ascontiguousarray(rho[:, ::-1, :])
    A2 = prc.Projector(rho2, (0.9, 1.2, 1.5), "18F", 16)
    x2 = np.ascontiguousarray(x[:, ::-1, :])
    rho3 = np.ascontiguousarray(rho[1:, 2:, 3:])  # uneven grid: 11 x 12 x 13
    A3 = prc.Projector(rho3, (1.3, 0.8, 1.0), "11C", 24)
    x3 = np.ascontiguousarray(x[1:, 2:, 3:])
    singles = [prc.correct(A1, x, 5, 0.1, 20.0, True), prc.correct(A2, x2, 5, 0.05, 20.0, True),
               prc.correct(A1, y, 5, 0.0, 20.0, True), prc.correct(A3, x3, 5, 0.2, 20.0, True)]
    batch = prc.batch_correct([A1, A2, A1, A3], [x, x2, y, x3], 5, beta=[0.1, 0.05, 0.0, 0.2], gamma=20.0,
                              accelerate=True)
    for a, b in zip(batch, singles):
        np.testing.assert_array_equal(a, b)


def test_renormalise_reset_and_rerun_reproduce(prc, phantom):
    # The timed paths rebuild the weights and restart from resident data: results must repeat.
    rho, x, y, sup = phantom
    announce("normalise again / solver reset / batch re-run")
    A = prc.Projector(rho, SPACING, "68Ga", 32)
    W0, ax0 = A.W, A.forward(x)
    A.normalise()
    np.testing.assert_array_equal(A.W, W0)
    np.testing.assert_array_equal(A.forward(x), ax0)
    s = prc.Solver(A, x, 0.1, 20.0, True)
    first = s.iterate(4).estimate()
    second = s.reset().iterate(4).estimate()
    np.testing.assert_array_equal(first, second)
    b = prc.Batch([A, A], [x, y], beta=0.1, gamma=20.0, accelerate=True)
    b.run(4)
    r1 = [b.estimate(0), b.estimate(1)]
    A.normalise()
    b.run(4)
    np.testing.assert_array_equal(b.estimate(0), r1[0])
    np.testing.assert_array_equal(b.estimate(1), r1[1])
    np.testing.assert_array_equal(r1[0], first)
    b.run(3)  # another iteration count captures another graph
    np.testing.assert_array_equal(b.estimate(0), s.reset().iterate(3).estimate())


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
@pytest.mark.parametrize("memo", [True, False])
@pytest.mark.parametrize("exact", [False, True])
def test_rebuild_matches_first_build(prc, phantom, precision, memo, exact):
    # A projector's first build also counts the march statistics; later builds run the march
    # kernels without the counting (the near march then reads densities without masks): every
    # stored weight, and so W and both operators, must come out bit for bit the same.
    rho, x, y, sup = phantom
    announce(f"rebuild vs first build {precision} memo={memo} exact={exact}")
    A = prc.Projector(rho, SPACING, "68Ga", 32, precision=precision, near_memo=memo, exact_range=exact)
    W0, ax0, aty0, st0 = A.W, A.forward(x), A.adjoint(y), A.stats
    A.normalise()
    np.testing.assert_array_equal(A.W, W0)
    np.testing.assert_array_equal(A.forward(x), ax0)
    np.testing.assert_array_equal(A.adjoint(y), aty0)
    st1 = A.stats
    for k in ("vrs_near", "vrs_far", "deposits", "weight_entries", "target_entries"):
        assert st1[k] == st0[k], k
    A.close()


def test_profiled_run_matches_graph_run(prc, phantom):
    rho, x, y, sup = phantom
    A = prc.Projector(rho, SPACING, "68Ga", 32, precision="mixed")
    b = prc.Batch([A], [x], beta=0.03, gamma=20.0, accelerate=True)
    ref = b.run(5).estimate(0)
    prc.profile_enable(True)
    try:
        out = b.run(5).estimate(0)
        rep = prc.profile_report()
    finally:
        prc.profile_enable(False)
    np.testing.assert_array_equal(out, ref)
    assert rep["forward"]["count"] == 5 and rep["adjoint"]["count"] == 5
    assert all(v["mean_ms"] >= 0 for v in rep.values())


@pytest.fixture(scope="module")
def blocks():
    """Piecewise-constant 24x20x18 phantom (lung, soft tissue and bone slabs, an air corner and
    a sub-threshold pocket): most sources see one density over their 5x5x5 block."""
    rho = np.full((18, 20, 24), 1.0, np.float32)
    rho[:, :, :7] = 0.3
    rho[:, 13:, 15:] = 1.85
    rho[:4, :5, :5] = 1e-3
    rho[9:12, 8:11, 10:13] = 0.02
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 2, rho.shape).astype(np.float32)
    y = rng.uniform(0, 2, rho.shape).astype(np.float32)
    return rho, x, y


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_uniform_neighbourhood_memo_is_bit_identical(prc, blocks, precision):
    rho, x, y = blocks
    announce(f"near memo vs full march 24x20x18 {precision}")
    A = prc.Projector(rho, SPACING, "68Ga", 32, precision=precision)
    B = prc.Projector(rho, SPACING, "68Ga", 32, precision=precision, near_memo=False)
    sa, sb = A.stats, B.stats
    assert sa["near_uniform"] > sa["n_active"] // 4 and sb["near_uniform"] == 0
    assert sa["vrs_near"] < sb["vrs_near"] and sa["vrs_far"] == sb["vrs_far"]
    np.testing.assert_array_equal(A.W, B.W)
    np.testing.assert_array_equal(A.forward(x), B.forward(x))
    np.testing.assert_array_equal(A.adjoint(y), B.adjoint(y))
    A.close()
    B.close()


def test_uniform_neighbourhood_memo_matches_oracle(prc, oracle_mod, blocks):
    rho, x, y = blocks
    A = prc.Projector(rho, SPACING, "68Ga", 32)
    table, _ = A.table
    rho_p = np.maximum(rho, np.float32(1e-3))
    O = oracle_mod.Projector(rho_p, SPACING, "68Ga", 32, table=table)
    w_ref = np.zeros(rho.size)
    w_ref[O.active] = O.W
    np.testing.assert_allclose(A.W.ravel(), w_ref, rtol=2e-6, atol=0)
    assert rel_l2(A.forward(x), O.forward(x)) <= 1e-6
    assert rel_l2(A.adjoint(y), O.adjoint(y)) <= 1e-6
    A.close()


@pytest.fixture(scope="module")
def wide_rows():
    """150x6x5 phantom whose rows span up to three 64-voxel segments: every row has its own
    active range (ragged last segments, segment boundaries at different x from row to row), some
    rows hold sub-threshold gaps and one row holds no active voxel at all."""
    rng = np.random.default_rng(11)
    rho = rng.uniform(0.2, 1.9, (5, 6, 150)).astype(np.float32)
    for z in range(5):
        for y in range(6):
            a, b = int(rng.integers(0, 40)), int(rng.integers(100, 151))
            rho[z, y, :a] = 0.01
            rho[z, y, b:] = 0.01
    rho[1::2, :, 70:75] = 0.02
    rho[2, 3, :] = 0.01
    x = rng.uniform(0, 2, rho.shape).astype(np.float32)
    y = rng.uniform(0, 2, rho.shape).astype(np.float32)
    return rho, x, y


def test_multi_segment_rows_match_oracle(prc, oracle_mod, wide_rows):
    rho, x, y = wide_rows
    announce("multi-segment rows 150x6x5: FP32 vs oracle, mixed vs FP32")
    A = prc.Projector(rho, SPACING, "68Ga", 32)
    table, _ = A.table
    rho_p = np.maximum(rho, np.float32(1e-3))
    sup = rho_p >= np.float32(0.05)
    assert not sup[2, 3].any() and sup.sum(axis=2).max() > 128
    O = oracle_mod.Projector(rho_p, SPACING, "68Ga", 32, table=table)
    w_ref = np.zeros(rho.size)
    w_ref[O.active] = O.W
    np.testing.assert_allclose(A.W.ravel(), w_ref, rtol=2e-6, atol=0)
    ax, aty = A.forward(x), A.adjoint(y)
    assert rel_l2(ax, O.forward(x)) <= 1e-6
    assert rel_l2(aty, O.adjoint(y)) <= 1e-6
    # The FP16 copies use a different lane mapping (8 entries per lane, not 4) over the same
    # segments; they differ from FP32 by the weights' rounding only, and keep the identities.
    M = prc.Projector(rho, SPACING, "68Ga", 32, precision="mixed")
    mx, mty = M.forward(x), M.adjoint(y)
    assert rel_l2(mx, ax) <= 1e-3 and rel_l2(mty, aty) <= 1e-3
    for fx, fty in ((ax, aty), (mx, mty)):
        lhs = np.sum(fx.astype(np.float64) * y)
        rhs = np.sum(x.astype(np.float64) * fty)
        assert abs(lhs / rhs - 1) <= 1e-6
        assert abs(fx.sum(dtype=np.float64) / x[sup].sum(dtype=np.float64) - 1) <= 1e-6
        np.testing.assert_array_equal(fx[~sup], 0)
    A.close()
    M.close()


@pytest.mark.parametrize("iso,rays,spacing,air_cap,support", [
    ("68Ga", 2, SPACING, 50.0, 0.05),           # the fewest rays the CLI accepts
    ("68Ga", 130, SPACING, 50.0, 0.05),         # more than two 64-ray chunks of the shift-list builder
    ("82Rb", 32, (2.0, 2.0, 2.0), 50.0, 0.05),  # isotropic voxels, the longest-range isotope
    ("11C", 24, (0.5, 3.0, 1.0), 50.0, 0.05),   # strongly anisotropic voxels, a short-range isotope
    ("82Rb", 32, SPACING, 3.0, 0.05),           # rays die after about two voxels of air
    ("68Ga", 32, SPACING, 10.0, 0.5),           # a higher support threshold: less tissue, more air
])
def test_geometry_edges_match_oracle(prc, oracle_mod, phantom, iso, rays, spacing, air_cap, support):
    rho, x, y, _ = phantom
    announce(f"{iso} rays={rays} spacing={spacing} air_cap={air_cap} support={support}: GPU vs oracle 16x14x12")
    A = prc.Projector(rho, spacing, iso, rays, support_rho=support, air_cap_mm=air_cap)
    table, _ = A.table
    O = oracle_mod.Projector(rho, spacing, iso, rays, support_rho=support, air_cap=air_cap, table=table)
    w_ref = np.zeros(rho.size)
    w_ref[O.active] = O.W
    np.testing.assert_allclose(A.W.ravel(), w_ref, rtol=2e-6, atol=0)
    ax, aty = A.forward(x), A.adjoint(y)
    assert rel_l2(ax, O.forward(x)) <= 1e-6
    assert rel_l2(aty, O.adjoint(y)) <= 1e-6
    # The mixed form on the same geometry: FP16 weight rounding only, identities kept.
    M = prc.Projector(rho, spacing, iso, rays, support_rho=support, air_cap_mm=air_cap, precision="mixed")
    mx, mty = M.forward(x), M.adjoint(y)
    assert rel_l2(mx, ax) <= 1e-3 and rel_l2(mty, aty) <= 1e-3
    lhs = np.sum(mx.astype(np.float64) * y)
    rhs = np.sum(x.astype(np.float64) * mty)
    assert abs(lhs / rhs - 1) <= 1e-6
    A.close()
    M.close()


@pytest.mark.parametrize("iso", ["68Ga", "124I"])
def test_exact_range_matches_oracle(prc, oracle_mod, phantom, iso):
    # --exact-range: the survival model evaluated at every step (prc.cu:75, 625) on both sides.
    # Device and host transcendentals may differ by an ulp, and a deposit is a difference of two
    # survivals: the bar is the per-iteration one, 1e-5.
    rho, x, y, _ = phantom
    announce(f"exact range {iso}: GPU vs oracle 16x14x12")
    A = prc.Projector(rho, SPACING, iso, 32, exact_range=True)
    O = oracle_mod.Projector(rho, SPACING, iso, 32, exact=True)
    w_ref = np.zeros(rho.size)
    w_ref[O.active] = O.W
    np.testing.assert_allclose(A.W.ravel(), w_ref, rtol=1e-5, atol=0)
    ef, ea = rel_l2(A.forward(x), O.forward(x)), rel_l2(A.adjoint(y), O.adjoint(y))
    assert ef <= 1e-5 and ea <= 1e-5, (ef, ea)
    A.close()


def test_mixed_batch_of_wide_rows_equals_single_runs(prc, wide_rows, phantom):
    # The batched mixed form over volumes of one and of several segments per row.
    rho, x, _ = wide_rows
    rho2, x2, _, _ = phantom
    announce("mixed batch: 150x6x5 and 16x14x12 vs single runs")
    A = prc.Projector(rho, SPACING, "68Ga", 32, precision="mixed")
    B = prc.Projector(rho2, SPACING, "18F", 16, precision="mixed")
    singles = [prc.correct(A, x, 4, 0.03, 20.0, True), prc.correct(B, x2, 4, 0.1, 20.0, True)]
    batch = prc.batch_correct([A, B], [x, x2], 4, beta=[0.03, 0.1], gamma=20.0, accelerate=True)
    for a, b in zip(batch, singles):
        np.testing.assert_array_equal(a, b)
    A.close()
    B.close()


def test_concurrent_normalisation_matches_single(prc, phantom, blocks):
    # Several projectors normalised together (one stream each) give the single builds' weights.
    rho, x, y, sup = phantom
    rho2, x2, _ = blocks
    announce("concurrent normalisation of three projectors")
    P = [prc.Projector(rho, SPACING, "68Ga", 32), prc.Projector(rho2, SPACING, "18F", 16, precision="mixed"),
         prc.Projector(rho, (0.9, 1.2, 1.5), "124I", 24)]
    ref = [(p.W, p.forward(v), p.adjoint(v)) for p, v in zip(P, (x, x2, y))]
    prc.normalise_all(P)
    for p, v, (w, f, a) in zip(P, (x, x2, y), ref):
        np.testing.assert_array_equal(p.W, w)
        np.testing.assert_array_equal(p.forward(v), f)
        np.testing.assert_array_equal(p.adjoint(v), a)
    for p in P:
        p.close()


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_every_schedule_gives_the_same_correction(prc, phantom, blocks, precision, monkeypatch):
    # The projector passes compute each item the same way under every launch shape (launch order,
    # XCD-local queues of 1, 2 or 4 blocks per CU, XCD-chunked launch order; gathers with or without
    # the sc0 cache policy): the batch result must not depend on it.
    rho, x, y, sup = phantom
    rho2, x2, _ = blocks
    announce(f"projector schedules ({precision})")
    A1 = prc.Projector(rho, SPACING, "68Ga", 32, precision=precision)
    A2 = prc.Projector(rho2, SPACING, "18F", 16, precision=precision)
    results = {}
    shapes = [n + pol for pol in ("", "+sc0") for n in ("grid", "chunked", "queue@1", "queue@2", "queue@4")]
    for sched in shapes:
        monkeypatch.setenv("PRC_TUNE", sched)
        results[sched] = prc.batch_correct([A1, A2, A1], [x, x2, y], 4, beta=0.1, gamma=20.0, accelerate=True)
    for sched in shapes[1:]:
        for a, b in zip(results[sched], results["grid"]):
            np.testing.assert_array_equal(a, b, err_msg=sched)
    A1.close()
    A2.close()


def test_mixed_form_is_scale_equivariant(prc, phantom):
    # The mixed solver keeps FP16 copies of its forward input; a per-volume power-of-two scale
    # keeps them in range, so data 2^13 times larger (values far beyond the FP16 maximum) give
    # exactly 2^13 times the correction (MLEM with extrapolation is scale-equivariant).
    rho, x, y, sup = phantom
    announce("mixed form: data scaled by 2^13")
    A = prc.Projector(rho, SPACING, "68Ga", 32, precision="mixed")
    d = np.where(sup, x * np.float32(64.0), 0).astype(np.float32)
    r1 = prc.correct(A, d, 4, 0.0, 20.0, True)
    r2 = prc.correct(A, d * np.float32(8192.0), 4, 0.0, 20.0, True)
    assert np.isfinite(r2).all() and float(d.max()) * 8192 > 65504
    np.testing.assert_array_equal(r2, r1 * np.float32(8192.0))
    A.close()


def test_mixed_form_keeps_a_wide_dynamic_range(prc):
    """Activity spanning seven decades (background 1e-3 on the 64^3 water cylinder, three voxels
    at 1e4): the FP16 copies that the mixed passes gather are scaled per iteration by their own
    maximum, so the background stays in normal FP16 numbers. Mixed against FP32 within 1e-3 overall
    and on the background more than 5 voxels from the hot voxels."""
    rho, _ = phantoms.cylinder_phantom()
    sup = np.maximum(rho, np.float32(1e-3)) >= np.float32(0.05)
    g = np.where(sup, np.float32(1e-3), np.float32(0)).astype(np.float32)
    hot = [(32, 32, 32), (20, 40, 30), (44, 25, 36)]
    for v in hot:
        g[v] = 1e4
    announce("mixed form, 1e7 dynamic range, 64^3 cylinder")
    A = prc.Projector(rho, (2.0, 2.0, 2.0), "68Ga", 32)
    M = prc.Projector(rho, (2.0, 2.0, 2.0), "68Ga", 32, precision="mixed")
    ref = prc.correct(A, g, 6, 0.03, 20.0, True)
    out = prc.correct(M, g, 6, 0.03, 20.0, True)
    assert rel_l2(out, ref) <= 1e-3
    far = sup.copy()
    for v in hot:
        far[v[0] - 5:v[0] + 6, v[1] - 5:v[1] + 6, v[2] - 5:v[2] + 6] = False
    assert far.sum() > 10000
    assert rel_l2(out[far], ref[far]) <= 1e-3


def test_device_ray_set_matches_closed_forms_and_oracle(prc, oracle_mod):
    """The device's ray set (prc.cu:118-129): against the FP64 closed forms of the multipliers at
    pair indices up to 2^20, where a one-digit change of any multiplier moves offsets or azimuths by
    more than 1e-4, and against the oracle's FP32 rays."""
    from test_oracle import closed_form_rays
    for n in (2, 64, 2048, 1 << 21):
        announce(f"device ray set n={n}")
        start, direction = prc.ray_set(n)
        ks = np.unique(np.concatenate([np.arange(min(n, 8)), np.random.default_rng(n).integers(0, n, 2000)]))
        want_s, want_d = closed_form_rays(ks, n)
        np.testing.assert_allclose(start[ks], want_s, rtol=0, atol=1e-7)
        np.testing.assert_allclose(direction[ks], want_d, rtol=0, atol=4e-7)
        o = [oracle_mod.ray_f32(int(k), n) for k in ks[:200]]
        np.testing.assert_allclose(start[ks[:200]], np.array([v[0] for v in o]), rtol=0, atol=6e-8)
        np.testing.assert_allclose(direction[ks[:200]], np.array([v[1] for v in o]), rtol=0, atol=6e-8)


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_far_marched_projector_matches_stored_and_oracle(prc, oracle_mod, pair, phantom, precision):
    """Near block stored only, far rays marched in every pass (the memory fallback): same W bit for
    bit as the stored form, operators and per-iteration corrections against the oracle (FP32) or the
    stored mixed form, the reference's invariants."""
    rho, x, y, sup = phantom
    A, O = pair
    announce(f"far-marched projector ({precision})")
    S = A if precision == "fp32" else prc.Projector(rho, SPACING, "68Ga", 32, precision="mixed")
    M = prc.Projector(rho, SPACING, "68Ga", 32, precision=precision, march_passes=True)
    st = M.stats
    assert st["far_marched"] == 1 and S.stats["far_marched"] == 0
    assert st["weight_entries"] < S.stats["weight_entries"]
    np.testing.assert_array_equal(M.W, S.W)
    ax, aty = M.forward(x), M.adjoint(y)
    if precision == "fp32":
        assert rel_l2(ax, O.forward(x)) <= 1e-6
        assert rel_l2(aty, O.adjoint(y)) <= 1e-6
    else:
        assert rel_l2(ax, S.forward(x)) <= 1e-6
        assert rel_l2(aty, S.adjoint(y)) <= 1e-6
    lhs, rhs = np.sum(ax.astype(np.float64) * y), np.sum(x.astype(np.float64) * aty)
    assert abs(lhs / rhs - 1) <= 1e-6
    assert abs(ax.sum(dtype=np.float64) / x.sum(dtype=np.float64) - 1) <= 1e-6
    assert not np.any(ax[~sup])
    if precision == "fp32":
        _, hist = O.solve(x, 6, 0.1, 20.0, True, history=True)
        s = prc.Solver(M, x, 0.1, 20.0, True)
        for it, ref in enumerate(hist, 1):
            s.iterate(1)
            assert rel_l2(s.estimate(), ref) <= 1e-5, f"iteration {it}"
        s.close()
    else:
        # The mixed solver rounds its gathered iterates to FP16 every iteration, so operator
        # differences at float rounding (the forward's atomic order) move some of those roundings:
        # the bar is the mixed form's own, against the stored mixed form and the FP32 oracle.
        out = prc.correct(M, x, 6, 0.1, 20.0, True)
        assert rel_l2(out, prc.correct(S, x, 6, 0.1, 20.0, True)) <= 1e-3
        assert rel_l2(out, O.solve(x, 6, 0.1, 20.0, True)) <= 1e-3
    # A batch mixing far-marched and stored volumes: each volume as its single run, to rounding.
    outs = prc.batch_correct([M, S], [x, y], 4, 0.1, 20.0, True)
    assert rel_l2(outs[0], prc.correct(M, x, 4, 0.1, 20.0, True)) <= (1e-6 if precision == "fp32" else 1e-4)
    np.testing.assert_array_equal(outs[1], prc.correct(S, y, 4, 0.1, 20.0, True))
    M.close()


def test_build_falls_back_to_marching_when_the_matrix_does_not_fit(prc, pair, phantom, monkeypatch):
    """A device-memory budget smaller than the stored matrix: the build switches to the far-marched
    form by itself and gives the stored form's operator."""
    rho, x, y, sup = phantom
    A, _ = pair
    monkeypatch.setenv("PRC_WEIGHT_BUDGET", "1000000")  # bytes
    announce("build under a 1 MB weight budget")
    M = prc.Projector(rho, SPACING, "68Ga", 32)
    assert M.stats["far_marched"] == 1
    np.testing.assert_array_equal(M.W, A.W)
    assert rel_l2(M.forward(x), A.forward(x)) <= 1e-6
    assert rel_l2(M.adjoint(y), A.adjoint(y)) <= 1e-6
    M.normalise()  # a rebuild keeps the form
    assert M.stats["far_marched"] == 1
    np.testing.assert_array_equal(M.W, A.W)
    M.close()


def test_nonzero_entry_statistic(pair):
    A, _ = pair
    st = A.stats
    assert st["deposits"] < st["nonzero_entries"] <= st["weight_entries"]
    assert st["nonzero_entries"] <= st["deposits"] + 125 * st["n_active"]
