"""GPU parity of the spectral (S) path: 8-tap periodic stencil, set-up,
incident field, right-hand side, warm start and solve_frequency against the
oracle (oracle/spectral.py) and the reference's golden vectors."""
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import spectral as S

pytestmark = pytest.mark.gpu
G = np.load(Path(__file__).parent / "golden" / "spectral_small.npz")


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def golden_spec():
    return dict(medium_shape=[int(G["h"]), int(G["w"])], dx_m=float(G["dx"]), x0_m=float(G["x0"]),
                z0_m=float(G["z0"]), angles_deg=list(G["angles"]), source_tref_s=list(G["trefs"]))


def golden_model(iterations=4, **kw):
    from nsrno_amd.wave_operator import WaveResidualOperator
    return WaveResidualOperator(dict(golden_spec(), **kw), iterations=iterations)


def medium_tensor(c, rho):
    return torch.tensor(np.stack((c, rho))[None], dtype=torch.float32)


def oracle_setup(c=None, rho=None, **kw):
    box = S.Box(int(G["h"]), int(G["w"]), float(G["dx"]), float(G["x0"]), float(G["z0"]))
    c = G["c"] if c is None else c
    rho = G["rho"] if rho is None else rho
    # the device receives float32 c and rho, like the reference (medium.float())
    return box, S.setup(box, np.float32(c), np.float32(rho), float(G["f"]), **kw)


def test_stag8_apply_fp64_fp32_batched():
    model = golden_model()
    data, _, _ = model.setup(medium_tensor(G["c"], G["rho"]), float(G["f"]))
    box, (ax, az, kc, k0) = oracle_setup()
    rng = np.random.default_rng(11)
    u = rng.normal(size=(3, box.H, box.W)) + 1j * rng.normal(size=(3, box.H, box.W))
    ref = S.apply_stencil(u, ax, az, kc, k0, box.dx)
    out64 = model.apply_operator(torch.from_numpy(u).cuda(), data).cpu().numpy()
    assert rel(out64, ref) < 1e-12
    out32 = model.apply_operator(torch.from_numpy(u.astype(np.complex64)).cuda(), data).cpu().numpy()
    assert rel(out32, ref) < 5e-6
    # against the reference's own complex64 FFT operator on its seeded field
    lu = model.apply_operator(torch.from_numpy(G["u"]).cuda(), data).cpu().numpy()
    assert rel(lu, G["lu"]) < 5e-6
    # the library's forward operator is k0^2 conj(L)
    fwd_ref = k0 ** 2 * np.conj(S.apply_stencil(np.conj(u), ax, az, kc, k0, box.dx))
    fwd = data.plan.apply(torch.from_numpy(u).cuda()).cpu().numpy()
    assert rel(fwd, fwd_ref) < 1e-12
    # in FP32 both directions (the adjoint takes the conjugate mass through the forward kernel)
    fwd32 = data.plan.apply(torch.from_numpy(u.astype(np.complex64)).cuda()).cpu().numpy()
    assert rel(fwd32, fwd_ref) < 5e-6
    adj32 = data.plan.apply(torch.from_numpy(u.astype(np.complex64)).cuda(), adjoint=True).cpu().numpy()
    assert rel(adj32, k0 ** 2 * ref) < 5e-6


@pytest.mark.parametrize("attenuation", [None, "relaxation"])
def test_setup_coefficients(attenuation):
    kw = {} if attenuation is None else dict(attenuation="relaxation", eta_per_unit_alpha=list(G["relax_epa"]),
                                              taus_s=list(G["relax_taus"]))
    model = golden_model(**kw)
    data, _, _ = model.setup(medium_tensor(G["tissue_c"], G["tissue_rho"]), float(G["f"]))
    okw = {} if attenuation is None else dict(attenuation="relaxation", eta_per_unit_alpha=G["relax_epa"],
                                               taus=G["relax_taus"])
    box, (ax, az, kc, k0) = oracle_setup(G["tissue_c"], G["tissue_rho"], **okw)
    assert abs(data.k0 - k0) < 1e-12 * k0
    kap = model.coefficients(data).cpu().numpy()  # L 1 = kappa_S
    assert rel(kap, kc) < 1e-12
    gold = G["power_kc_tissue"] if attenuation is None else G["relax_kc"]
    assert rel(kap, gold) < 2e-6


def test_incident_and_rhs():
    model = golden_model()
    f = float(G["f"])
    data, _, _ = model.setup(medium_tensor(G["c"], G["rho"]), f)
    box, (ax, az, kc, k0) = oracle_setup()
    for a in range(3):
        theta = float(G["angles"][a]) * np.pi / 180
        ref, _, _ = S.incident(box, f, theta)
        inc = model.incident64(f, a, data)
        assert rel(inc.cpu().numpy(), ref) < 1e-11, a
    assert rel(model.incident(f, 1, data).cpu().numpy(), G["inc"]) < 5e-5
    ref, _, _ = S.incident(box, f, float(G["angles"][1]) * np.pi / 180)
    rhs = model.scattered_rhs(torch.from_numpy(ref).cuda(), data).cpu().numpy()
    assert rel(rhs, k0 ** 2 * S.scattered_rhs(box, ref, ax, az, kc, k0)) < 1e-12


def test_source_factor():
    model = golden_model()
    sf = complex(model.source_factor(float(G["f"]), 2))
    assert abs(sf - complex(G["sf"])) < 2e-6 * abs(sf)


def test_warm_start_never_increases_residual():
    model = golden_model()
    f = float(G["f"])
    data, _, _ = model.setup(medium_tensor(G["c"], G["rho"]), f)
    rhs = model.scattered_rhs(model.incident64(f, 0, data)[None], data)
    x = torch.zeros_like(rhs)
    res = [1.0]
    for _ in range(6):
        data.plan.warm_start(rhs, x, 1, adjoint=True)
        r = rhs - data.plan.apply(x, adjoint=True)
        res.append(float(r.norm() / rhs.norm()))
    assert all(b <= a * (1 + 1e-6) for a, b in zip(res, res[1:])), res
    assert res[-1] < 0.9


def test_solve_frequency_matches_oracle_and_reference():
    from nsrno_amd.wave_solver import solve_frequency
    model = golden_model()
    f = float(G["f"])
    pred, info = solve_frequency(model, medium_tensor(G["c"], G["rho"]), f, 1, tolerance=1e-9,
                                 max_iterations=3200, restart=32)
    assert set(info) == {"reason", "initial_neural_iterations", "solver_iterations", "relative_residual", "seconds",
                         "trace"}
    assert info["reason"] == "converged" and info["relative_residual"] < 1e-9
    assert info["initial_neural_iterations"] == 4
    assert pred.dtype == torch.complex64 and pred.shape == (192,)
    box = S.Box(int(G["h"]), int(G["w"]), float(G["dx"]), float(G["x0"]), float(G["z0"]))
    theta = float(G["angles"][1]) * np.pi / 180
    opred, ou, oinfo = S.solve_gmres(box, np.float32(G["c"]), np.float32(G["rho"]), f, theta, float(G["trefs"][1]),
                                     iterations=4, tolerance=1e-11)
    assert rel(pred.cpu().numpy().astype(np.complex128), opred) < 1e-5
    # the reference's complex64 solve stagnated at 8e-6 (iteration_budget)
    assert rel(pred.cpu().numpy(), G["pred"]) < 0.03


def test_solve_frequency_batch_fields():
    from nsrno_amd.wave_solver import solve_frequency_batch
    model = golden_model()
    f = float(G["f"])
    preds, x, infos = solve_frequency_batch(model, medium_tensor(G["c"], G["rho"]), f, [0, 1, 2], tolerance=1e-9,
                                            max_iterations=3200, restart=32)
    box = S.Box(int(G["h"]), int(G["w"]), float(G["dx"]), float(G["x0"]), float(G["z0"]))
    for k in range(3):
        assert infos[k]["reason"] == "converged"
        theta = float(G["angles"][k]) * np.pi / 180
        opred, ou, _ = S.solve_gmres(box, np.float32(G["c"]), np.float32(G["rho"]), f, theta,
                                     float(G["trefs"][k]), iterations=4, tolerance=1e-11)
        assert rel(x[k].cpu().numpy(), ou) < 1e-5, k
        assert rel(preds[k].cpu().numpy().astype(np.complex128), opred) < 1e-5, k


def test_solve_frequency_guards():
    from nsrno_amd.wave_solver import solve_frequency
    model = golden_model()
    m = medium_tensor(G["c"], G["rho"])
    with pytest.raises(ValueError):
        solve_frequency(model, m.clone().requires_grad_(True), float(G["f"]), 0)
    with pytest.raises(ValueError):
        solve_frequency(model, m, float(G["f"]), 0, max_iterations=100, restart=32)
    pred, info = solve_frequency(model, m, float(G["f"]), 0, tolerance=1e-12, max_iterations=32, restart=32)
    assert info["reason"] == "iteration_budget"
