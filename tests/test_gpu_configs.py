"""The five benchmark workloads at their real sizes on the GPU, each against the oracle or the
invariants the reference's tests hold (tests/test_solver.py:180-190, 245-255, 304-333).

C1: 64^3 at 2 mm, 18F water cylinder, 64 rays, Richardson-Lucy without acceleration, 10 iterations.
C2: 128x128x159 at 0.776x0.776x0.795 mm, 68Ga thorax, 128 rays, beta 0.03, gamma 20, accelerated,
    6 iterations.
C3: the C1 cylinder with 82Rb and 124I, 256 rays, 20 iterations, mixed form.
C4: 12 uneven volumes (96..160 x 96..160 x 80..200 voxels of 0.5-1.2 mm, 18F/11C/64Cu/68Ga,
    beta 0-0.2) batched, FP32 and mixed.
C5: 4 clinical-scale thoraxes (344x344x127 at 2.04x2.04x2.03 mm), 68Ga, 256 rays, mixed batched,
    20 iterations.
Data are the truth blurred by the forward projector and Poisson-noised (seeded)."""
import numpy as np
import pytest

from conftest import announce, rel_l2
from prc_mlem_amd import phantoms

pytestmark = pytest.mark.gpu


def noisy_data(prc, rho, truth, spacing, iso, rays, counts, seed, precision="fp32"):
    A = prc.Projector(rho, spacing, iso, rays, precision=precision)
    blurred = A.forward(truth)
    A.close()
    return phantoms.poisson_counts(blurred, counts, seed)


def identity_and_counts(A, sup, seed, bound=1e-6):
    """<Ax, y> = <x, A^T y>, sum(Ax) = sum(x), A^T 1 = 1 on the support, Ax = 0 off it."""
    rng = np.random.default_rng(seed)
    x = (rng.uniform(0.1, 2.0, sup.shape) * sup).astype(np.float32)
    y = (rng.uniform(0.1, 2.0, sup.shape) * sup).astype(np.float32)
    ax, aty = A.forward(x), A.adjoint(y)
    lhs, rhs = np.sum(ax.astype(np.float64) * y), np.sum(x.astype(np.float64) * aty)
    assert abs(lhs / rhs - 1) <= bound
    assert abs(ax.sum(dtype=np.float64) / x.sum(dtype=np.float64) - 1) <= bound
    assert not np.any(ax[~sup])
    at1 = A.adjoint(np.ones(sup.shape, np.float32))
    np.testing.assert_allclose(at1[sup], 1, rtol=3e-6, atol=3e-6)


def per_iteration_parity(prc, oracle_mod, A, O, g, iterations, beta, gamma, accel, bound=1e-5):
    """GPU FP32 solver against the FP32 oracle after every iteration; returns the last estimate."""
    _, gp, sup, support_sum, _ = oracle_mod.preprocess(O.rho, g)
    _, hist = O.solve(gp, iterations, beta, gamma, accel, history=True)
    s = prc.Solver(A, g, beta, gamma, accel)
    try:
        for it, ref in enumerate(hist, 1):
            s.iterate(1)
            est = s.estimate()
            err = rel_l2(est, ref)
            assert err <= bound, f"iteration {it}: {err:.3g}"
            assert not np.any(est[~sup])
        tot, ssum = s.sums()
        assert ssum == pytest.approx(support_sum, rel=1e-12)
        return est, tot / ssum - 1
    finally:
        s.close()


@pytest.mark.timeout(600)
def test_c1_cylinder_rl_matches_oracle_per_iteration(prc, oracle_mod):
    rho, truth = phantoms.cylinder_phantom()
    sp = (2.0, 2.0, 2.0)
    announce("C1: 64^3 18F cylinder, 64 rays, RL, 10 iterations vs the FP32 oracle")
    g = noisy_data(prc, rho, truth, sp, "18F", 64, 1e6, seed=3001)
    A = prc.Projector(rho, sp, "18F", 64)
    rho_p = np.maximum(rho, np.float32(1e-3))
    O = oracle_mod.Projector(rho_p, sp, "18F", 64, table=A.table[0])
    W = A.W.ravel()
    np.testing.assert_allclose(W[O.active], O.W, rtol=2e-6, atol=0)
    est, count_err = per_iteration_parity(prc, oracle_mod, A, O, g, 10, 0.0, 20.0, False)
    assert abs(count_err) <= 1e-6  # Richardson-Lucy keeps the supported counts
    identity_and_counts(A, rho_p >= np.float32(0.05), 31)


@pytest.fixture(scope="module")
def thorax(prc):
    rho, truth = phantoms.thorax_phantom()
    sp = (0.776, 0.776, 0.795)
    announce("C2: blur and noise the 128x128x159 68Ga thorax")
    g = noisy_data(prc, rho, truth, sp, "68Ga", 128, 5e7, seed=3002)
    return rho, g, sp


@pytest.mark.timeout(900)
def test_c2_thorax_fp32_matches_oracle_per_iteration(prc, oracle_mod, thorax):
    rho, g, sp = thorax
    announce("C2: 128x128x159 68Ga, 128 rays, beta 0.03, accelerated, 6 iterations vs the FP32 oracle")
    A = prc.Projector(rho, sp, "68Ga", 128)
    rho_p = np.maximum(rho, np.float32(1e-3))
    O = oracle_mod.Projector(rho_p, sp, "68Ga", 128, table=A.table[0])
    np.testing.assert_allclose(A.W.ravel()[O.active], O.W, rtol=2e-6, atol=0)
    _, count_err = per_iteration_parity(prc, oracle_mod, A, O, g, 6, 0.03, 20.0, True)
    assert abs(count_err) <= 1e-6  # the prior's count normalisation keeps the supported counts
    identity_and_counts(A, rho_p >= np.float32(0.05), 32)


@pytest.mark.timeout(600)
def test_c2_thorax_mixed_against_fp32(prc, thorax):
    rho, g, sp = thorax
    announce("C2: mixed form against FP32 on the full thorax")
    A = prc.Projector(rho, sp, "68Ga", 128)
    M = prc.Projector(rho, sp, "68Ga", 128, precision="mixed")
    ref = prc.correct(A, g, 6, 0.03, 20.0, True)
    out = prc.correct(M, g, 6, 0.03, 20.0, True)
    assert rel_l2(out, ref) <= 1e-3
    assert abs(out.sum(dtype=np.float64) / ref.sum(dtype=np.float64) - 1) <= 1e-6
    identity_and_counts(M, np.maximum(rho, np.float32(1e-3)) >= np.float32(0.05), 33)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("iso", ["82Rb", "124I"])
def test_c3_long_range_isotopes(prc, oracle_mod, iso):
    rho, truth = phantoms.cylinder_phantom()
    sp = (2.0, 2.0, 2.0)
    announce(f"C3: 64^3 cylinder {iso}, 256 rays, 20 iterations: FP32 vs the oracle, mixed vs FP32")
    g = noisy_data(prc, rho, truth, sp, iso, 256, 1e6, seed=3003)
    A = prc.Projector(rho, sp, iso, 256)
    rho_p = np.maximum(rho, np.float32(1e-3))
    O = oracle_mod.Projector(rho_p, sp, iso, 256, table=A.table[0])
    ref32, count_err = per_iteration_parity(prc, oracle_mod, A, O, g, 20, 0.03, 20.0, True)
    assert abs(count_err) <= 1e-6
    M = prc.Projector(rho, sp, iso, 256, precision="mixed")
    out = prc.correct(M, g, 20, 0.03, 20.0, True)
    assert rel_l2(out, ref32) <= 1e-3
    identity_and_counts(M, rho_p >= np.float32(0.05), 34)
    # Against the FP64 oracle (exact range model at every step) on a central 24^3 crop.
    crop = (slice(20, 44),) * 3
    r, gc = np.ascontiguousarray(rho_p[crop]), np.ascontiguousarray(g[crop])
    _, gcp, _, _, _ = oracle_mod.preprocess(r, gc)
    ref64 = oracle_mod.Projector(r, sp, iso, 256, prec=64).solve(gcp, 20, 0.03, 20.0, True)
    Mc = prc.Projector(r, sp, iso, 256, precision="mixed")
    assert rel_l2(prc.correct(Mc, gc, 20, 0.03, 20.0, True), ref64) <= 1e-3


def batch_vs_single(prc, vols, datas, precision, iterations):
    projs = [prc.Projector(v["rho"], v["spacing"], v["isotope"], 128, precision=precision) for v in vols]
    betas = [v["beta"] for v in vols]
    outs = prc.batch_correct(projs, datas, iterations, betas, 20.0, True)
    for k, (A, v, g, out) in enumerate(zip(projs, vols, datas, outs)):
        single = prc.correct(A, g, iterations, v["beta"], 20.0, True)
        np.testing.assert_array_equal(out, single, err_msg=f"volume {k}")
        sup = np.maximum(v["rho"], np.float32(1e-3)) >= np.float32(0.05)
        ssum = np.where(sup, np.maximum(g, 0), 0).sum(dtype=np.float64)
        assert abs(out.sum(dtype=np.float64) / ssum - 1) <= 1e-6, f"volume {k}"
        assert not np.any(out[~sup]) and out.min() >= 0
    stored = sum(A.stats["weight_bytes"] for A in projs)
    for A in projs:
        A.close()
    return outs, stored


@pytest.mark.timeout(900)
def test_c4_uneven_batch_fp32_and_mixed(prc):
    vols = phantoms.uneven_batch()
    announce("C4: 12 uneven volumes, blur and noise")
    datas = [noisy_data(prc, v["rho"], v["truth"], v["spacing"], v["isotope"], 128, v["counts"], 3004 + k)
             for k, v in enumerate(vols)]
    announce("C4: FP32 batch of 12 against single runs")
    out32, bytes32 = batch_vs_single(prc, vols, datas, "fp32", 6)
    announce(f"C4: FP32 stored matrix {bytes32 / 2**30:.1f} GiB; mixed batch of 12 against single runs")
    outm, bytesm = batch_vs_single(prc, vols, datas, "mixed", 6)
    announce(f"C4: mixed stored matrix {bytesm / 2**30:.1f} GiB")
    for k, (a, b) in enumerate(zip(outm, out32)):
        assert rel_l2(a, b) <= 1e-3, f"volume {k}"


@pytest.mark.timeout(900)
def test_c5_clinical_thorax_mixed_batch(prc):
    sp = (2.04, 2.04, 2.03)
    rhos, datas = [], []
    for v in range(4):
        announce(f"C5: thorax {v}: 344x344x127, blur and noise")
        rho, truth = phantoms.clinical_thorax(seed=1005 + v)
        rhos.append(rho)
        datas.append(noisy_data(prc, rho, truth, sp, "68Ga", 256, 2e8, 3005 + v, precision="mixed"))
    projs = [prc.Projector(r, sp, "68Ga", 256, precision="mixed") for r in rhos]
    stored = sum(A.stats["weight_bytes"] for A in projs)
    announce(f"C5: mixed stored matrix of 4 volumes {stored / 2**30:.1f} GiB; batched, 20 iterations")
    outs = prc.batch_correct(projs, datas, 20, 0.03, 20.0, True)
    for k, (rho, g, out) in enumerate(zip(rhos, datas, outs)):
        sup = np.maximum(rho, np.float32(1e-3)) >= np.float32(0.05)
        assert np.isfinite(out).all() and out.min() >= 0 and not np.any(out[~sup])
        ssum = np.where(sup, g, 0).sum(dtype=np.float64)
        assert abs(out.sum(dtype=np.float64) / ssum - 1) <= 1e-6, f"volume {k}"
    single = prc.correct(projs[0], datas[0], 20, 0.03, 20.0, True)
    np.testing.assert_array_equal(outs[0], single)
    sup0 = np.maximum(rhos[0], np.float32(1e-3)) >= np.float32(0.05)
    identity_and_counts(projs[0], sup0, 35)
    for A in projs[1:]:
        A.close()
    announce("C5: FP32 single run of volume 0 against the mixed batch")
    A32 = prc.Projector(rhos[0], sp, "68Ga", 256)
    ref = prc.correct(A32, datas[0], 20, 0.03, 20.0, True)
    assert rel_l2(outs[0], ref) <= 1e-3


@pytest.mark.timeout(600)
def test_c2_far_marched_form_matches_stored(prc, thorax):
    """The memory fallback at the thorax's full size: W bit for bit, operators and a 6-iteration
    correction against the stored form."""
    rho, g, sp = thorax
    announce("C2: far rays marched in every pass, against the stored form")
    A = prc.Projector(rho, sp, "68Ga", 128)
    M = prc.Projector(rho, sp, "68Ga", 128, march_passes=True)
    assert M.stats["far_marched"] == 1
    assert M.stats["weight_bytes"] < A.stats["weight_bytes"] / 10
    np.testing.assert_array_equal(M.W, A.W)
    sup = np.maximum(rho, np.float32(1e-3)) >= np.float32(0.05)
    x = np.where(sup, g, 0).astype(np.float32)
    assert rel_l2(M.forward(x), A.forward(x)) <= 1e-6
    assert rel_l2(M.adjoint(x), A.adjoint(x)) <= 1e-6
    identity_and_counts(M, sup, 36)
    assert rel_l2(prc.correct(M, g, 6, 0.03, 20.0, True), prc.correct(A, g, 6, 0.03, 20.0, True)) <= 1e-5
