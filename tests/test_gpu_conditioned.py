"""GPU parity of the conditioned split-PML operator (op_kind HELM_OP_COND8)
against the reference's golden vectors (tests/golden/conditioned_small.npz,
nsrno/conditioned_wave.py:52-88, nsrno/dense_wave.py:59-82) and the oracle
(oracle/conditioned.py): forward and adjoint applies in FP64 and FP32, the
source term, the adjoint identity, and multigrid-preconditioned solves whose
true residual the oracle recomputes, with the reciprocity <g, u> = <lam, b>
of a forward/adjoint pair."""
import json
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import conditioned as oc

pytestmark = pytest.mark.gpu
GOLD = dict(np.load(Path(__file__).parent / "golden" / "conditioned_small.npz"))
SPEC = json.loads(str(GOLD["spec_json"]))


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def gpu_system(spec, c, rho, alpha, f, **kw):
    from nsrno_amd.conditioned import ConditionedFrequencySystem, FrequencyGeometry
    g = FrequencyGeometry(spec)
    medium = torch.tensor(np.stack((c, rho)), dtype=torch.float64, device="cuda")
    return ConditionedFrequencySystem(g, medium, torch.tensor(alpha, dtype=torch.float64, device="cuda"), f, **kw)


def golden_system(**kw):
    return gpu_system(SPEC, GOLD["c"], GOLD["rho"], GOLD["alpha"], float(GOLD["f"]), **kw)


def test_golden_apply_adjoint_source():
    from nsrno_amd.conditioned import adjoint_operator, projected_operator, source_rhs
    s = golden_system()
    p = torch.from_numpy(GOLD["p"]).cuda()
    assert rel(projected_operator(s, p).cpu().numpy(), GOLD["proj"]) < 1e-12
    assert rel(adjoint_operator(s, p).cpu().numpy(), GOLD["adj"]) < 1e-12
    # the reference's __call__ on a field that vanishes outside the medium
    pm = p * s.g.mask
    call = oc.System(oc.Geometry(SPEC), np.stack((GOLD["c"], GOLD["rho"])), GOLD["alpha"], float(GOLD["f"]))
    assert rel(s(pm).cpu().numpy(), call.apply_fft(pm.cpu().numpy())) < 1e-12
    # FP32 path
    out32 = projected_operator(s, p.to(torch.complex64)).cpu().numpy()
    assert rel(out32, GOLD["proj"]) < 2e-6
    adj32 = adjoint_operator(s, p.to(torch.complex64)).cpu().numpy()
    assert rel(adj32, GOLD["adj"]) < 2e-6
    src = source_rhs(s, torch.from_numpy(GOLD["line"]).cuda()).cpu().numpy()
    assert rel(src, GOLD["src"]) < 1e-13
    # box coefficients (crop + replicate) as the reference's
    np.testing.assert_array_equal(s.c.cpu().numpy(), GOLD["sc"])
    fixed = s.fixed_preconditioner()(torch.from_numpy(GOLD["r"]).cuda()).cpu().numpy()
    assert rel(fixed, GOLD["fixed"]) < 1e-12
    s.close()


def larger_case(seed=2, h=120, w=152, pml=20, f=4.2e6):
    rng = np.random.default_rng(seed)
    zz, xx = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    blob = np.exp(-((zz - h / 2) ** 2 + (xx - w / 3) ** 2) / (2 * 12.0 ** 2))
    c = 1540.0 + 45.0 * blob + rng.normal(scale=2.0, size=(h, w))
    rho = 1000.0 - 50.0 * blob + rng.normal(scale=2.0, size=(h, w))
    alpha = 0.5 + 0.3 * blob
    # source row 4 cells inside the PML's inner edge; the 192 receivers of the reference's array
    # (38.2 mm) do not fit this box, so tests inject at their own columns
    spec = dict(SPEC, medium_shape=[2, h, w], pml=pml, x0_m=-w * 25e-6, z0_m=-(pml + 4) * 50e-6,
                source_c_lookup_zx=[30, 20])
    return spec, c, rho, alpha, f


@pytest.mark.parametrize("nrhs", [1, 3])
def test_apply_matches_oracle_with_exterior_values(nrhs):
    spec, c, rho, alpha, f = larger_case()
    s = gpu_system(spec, c, rho, alpha, f)
    o = oc.System(oc.Geometry(spec), np.stack((c, rho)), alpha, f)
    rng = np.random.default_rng(7)
    g = o.g
    u = rng.normal(size=(nrhs, g.H, g.W)) + 1j * rng.normal(size=(nrhs, g.H, g.W))  # nonzero outside too
    ut = torch.from_numpy(u).cuda()
    fwd = s(ut).cpu().numpy()
    adj = s.adjoint(ut).cpu().numpy()
    for r in range(nrhs):
        assert rel(fwd[r], o.projected(u[r])) < 1e-12
        assert rel(adj[r], o.adjoint(u[r])) < 1e-12
    fwd32 = s(ut.to(torch.complex64)).cpu().numpy()
    assert rel(fwd32, fwd) < 2e-6
    s.close()


def test_adjoint_identity_on_gpu():
    spec, c, rho, alpha, f = larger_case(seed=4)
    s = gpu_system(spec, c, rho, alpha, f)
    m = s.g.mask
    gen = torch.Generator(device="cuda").manual_seed(0)
    x = torch.randn(s.g.H, s.g.W, dtype=torch.complex128, device="cuda", generator=gen) * m
    y = torch.randn(s.g.H, s.g.W, dtype=torch.complex128, device="cuda", generator=gen) * m
    left = complex((y.conj() * s(x)).sum())
    right = complex((s.adjoint(y).conj() * x).sum())
    assert abs(left - right) / max(abs(left), abs(right)) < 1e-12
    s.close()


@pytest.mark.parametrize("method", ["gmres", "bicgstab"])
def test_solve_true_residual_and_reciprocity(method):
    from nsrno_amd.conditioned import solve_system, source_rhs
    spec, c, rho, alpha, f = larger_case(seed=5)
    s = gpu_system(spec, c, rho, alpha, f, nrhs=2)
    o = oc.System(oc.Geometry(spec), np.stack((c, rho)), alpha, f)
    rng = np.random.default_rng(9)
    line = rng.normal(size=(2, spec["medium_shape"][2])) + 1j * rng.normal(size=(2, spec["medium_shape"][2]))
    b = source_rhs(s, torch.from_numpy(line).cuda())
    assert rel(b.cpu().numpy(), oc.source_rhs(o, line)) < 1e-13
    u, info = solve_system(s, b, tolerance=1e-9, mixed=True, method=method, max_cycles=400)
    assert info["reason"] == "converged", info
    assert info["max_angle_residual"] < 1e-9
    un = u.cpu().numpy()
    bn = b.cpu().numpy()
    for r in range(2):  # the oracle's own FP64 residual
        assert rel(o.projected(un[r]), bn[r]) < 2e-9
    # adjoint solve with receiver-injected right-hand side (dense_wave.py:244)
    g = s.g
    grad = torch.zeros_like(b)
    cols = torch.arange(g.px + 25, g.px + g.w - 25, 2, device="cuda")
    assert int(cols.max()) < g.W
    wts = torch.from_numpy(rng.normal(size=(2, len(cols))) + 1j * rng.normal(size=(2, len(cols)))).cuda()
    grad[..., g.face, cols] = wts
    lam, ainfo = solve_system(s, grad, adjoint=True, tolerance=1e-9, method=method, max_cycles=400)
    assert ainfo["reason"] == "converged", ainfo
    ln = lam.cpu().numpy()
    for r in range(2):
        assert rel(o.adjoint(ln[r]), grad[r].cpu().numpy()) < 2e-9
        lhs = np.vdot(grad[r].cpu().numpy(), un[r])  # <g, u>
        rhs = np.vdot(ln[r], bn[r])                  # <lam, b>
        assert abs(lhs - rhs) / abs(lhs) < 1e-7
    s.close()


def test_scattered_rhs_matches_oracle():
    from nsrno_amd.conditioned import scattered_rhs
    spec, c, rho, alpha, f = larger_case(seed=6)
    s = gpu_system(spec, c, rho, alpha, f)
    h, w = spec["medium_shape"][1:]
    ref = gpu_system(spec, np.full((h, w), 1540.0), np.full((h, w), 1000.0), np.full((h, w), .002), f)
    o = oc.System(oc.Geometry(spec), np.stack((c, rho)), alpha, f)
    o0 = oc.System(oc.Geometry(spec), np.stack((np.full((h, w), 1540.0), np.full((h, w), 1000.0))),
                   np.full((h, w), .002), f)
    rng = np.random.default_rng(3)
    g = o.g
    pref = (rng.normal(size=(1, g.H, g.W)) + 1j * rng.normal(size=(1, g.H, g.W))) * g.mask
    line = rng.normal(size=(1, w)) + 1j * rng.normal(size=(1, w))
    got = scattered_rhs(s, ref, torch.from_numpy(pref).cuda(), torch.from_numpy(line).cuda()).cpu().numpy()
    want = oc.source_rhs(o, line) - oc.source_rhs(o0, line) - o.apply(pref[0]) + o0.apply(pref[0])
    assert rel(got, want) < 1e-12
    s.close()
    ref.close()


def test_direct_solve_field_accuracy():
    """Fields of the GPU solve within 1e-8 of the oracle's complex128 sparse LU."""
    from nsrno_amd.conditioned import solve_system, source_rhs
    s = golden_system(nrhs=2)
    o = oc.System(oc.Geometry(SPEC), np.stack((GOLD["c"], GOLD["rho"])), GOLD["alpha"], float(GOLD["f"]))
    b = source_rhs(s, torch.from_numpy(GOLD["line"]).cuda())
    u, info = solve_system(s, b, tolerance=1e-11, max_cycles=400)
    assert info["reason"] == "converged", info
    ref = oc.solve_direct(o, b.cpu().numpy())
    assert rel(u.cpu().numpy(), ref) < 1e-8
    lam, ainfo = solve_system(s, b, adjoint=True, tolerance=1e-11, max_cycles=400)
    assert rel(lam.cpu().numpy(), oc.solve_direct(o, b.cpu().numpy(), adjoint=True)) < 1e-8
    s.close()


def test_material_gradient_matches_oracle_fd():
    """helm_conditioned_gradient = d/d(c, rho) of G = -Re sum_r <lam_r, M (A v_r - s_r)> at fixed
    v, lam (the step of dense_wave.py:250-255), against central differences of the oracle's G."""
    from nsrno_amd.conditioned import material_gradient
    s = golden_system(nrhs=2)
    spec = SPEC
    g = oc.Geometry(spec)
    rng = np.random.default_rng(21)
    v = (rng.normal(size=(2, g.H, g.W)) + 1j * rng.normal(size=(2, g.H, g.W)))
    lam = (rng.normal(size=(2, g.H, g.W)) + 1j * rng.normal(size=(2, g.H, g.W)))
    line = GOLD["line"]
    grad = material_gradient(s, torch.from_numpy(v).cuda(), torch.from_numpy(lam).cuda(),
                             torch.from_numpy(line).cuda()).cpu().numpy()

    def G(c, rho):
        o = oc.System(g, np.stack((c, rho)), GOLD["alpha"], float(GOLD["f"]))
        src = oc.source_rhs(o, line)
        return -sum(np.vdot(lam[r], g.mask * (o.apply(g.mask * v[r]) - src[r])).real for r in range(2))

    from scipy.ndimage import gaussian_filter
    for k in range(3):
        dc = gaussian_filter(rng.normal(size=GOLD["c"].shape), 2.0) * 20.0
        dr = gaussian_filter(rng.normal(size=GOLD["c"].shape), 2.0) * 20.0
        if k == 1:
            dr[:] = 0.0
        if k == 2:
            dc[:] = 0.0
        eps = 1e-4
        fd = (G(GOLD["c"] + eps * dc, GOLD["rho"] + eps * dr) - G(GOLD["c"] - eps * dc, GOLD["rho"] - eps * dr)) / (2 * eps)
        an = float(np.sum(grad[0] * dc) + np.sum(grad[1] * dr))
        assert abs(an - fd) / abs(fd) < 1e-6, (k, an, fd)
    s.close()


def test_implicit_channels_gradient_fd():
    """dense_wave.ImplicitChannels semantics: autograd through the converged forward solve,
    adjoint solve and material gradient, against central differences of converged solves."""
    from nsrno_amd.conditioned import DenseCase, FrequencyGeometry, implicit_channels
    h, w, dx = 64, 800, 50e-6
    spec = dict(SPEC, medium_shape=[2, h, w], dx_m=dx, x0_m=-0.0191 - 10 * dx, z0_m=-30 * dx, native_dt_s=2.5e-9,
                pml=24, source_c_lookup_zx=[30, 12])
    g = FrequencyGeometry(spec)
    assert int(g.rx_cols.max()) < g.px + w and int(g.rx_cols.min()) >= g.px
    rng = np.random.default_rng(31)
    zz, xx = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    blob = np.exp(-((zz - 40) ** 2 + (xx - 400) ** 2) / (2 * 8.0 ** 2))
    f = 5.5e6
    lines = np.stack([np.hanning(w) * np.exp(-2j * np.pi * f * (np.arange(w) - w / 2) * dx * np.sin(t) / 1540.0)
                      for t in (-0.1, 0.15)])
    case = DenseCase(g, f, torch.from_numpy(lines).cuda())
    med = torch.tensor(np.stack((1540.0 + 30.0 * blob, 1000.0 + 40.0 * blob)), dtype=torch.float64, device="cuda")
    alpha = torch.full((h, w), 0.5, dtype=torch.float64, device="cuda")
    wts = torch.from_numpy(rng.normal(size=(2, 192)) + 1j * rng.normal(size=(2, 192))).cuda()

    def J(m):
        ch = implicit_channels(m, alpha, case, tolerance=1e-11)
        return (wts.conj() * ch).real.sum()

    m = med.clone().requires_grad_(True)
    J(m).backward()
    grad = m.grad.detach()
    dm = torch.zeros_like(med)
    dm[0] = torch.from_numpy(20.0 * blob).cuda()
    dm[1] = torch.from_numpy(-15.0 * blob).cuda()
    eps = 1e-3
    with torch.no_grad():
        fd = float((J(med + eps * dm) - J(med - eps * dm)) / (2 * eps))
    an = float((grad * dm).sum())
    assert abs(an - fd) / abs(fd) < 1e-4, (an, fd)
    case.close()
