"""GPU: the batched solve over plans of different grid sizes (helm_batch_solve).

* Equality: plans of three grid sizes solved together give, plan by plan,
  the fields, iteration counts and true residuals of helm_solve on the same
  plan (default launch shapes, no autotuning): the same kernel bodies run on
  the same arrays in the same order, so the fields agree bit for bit.
* Exactness: the batched fields of one frequency against the oracle's
  complex128 sparse LU (1e-5).
* Uneven convergence: plans that stop at different cycles leave the
  segment tables one by one; the graph is captured again each time and the
  plans still running are unaffected (an easy plan and a hard plan together
  equal each alone).
* The batched sweep (run_sweep(mode="batched")) at config 3's lowest and
  highest frequency: 11 angles, interior 384^2 and 854^2 + 32 PML, converged
  to 1e-6, traces equal to the threaded sweep's to the solve tolerance.
* Argument errors raise before any launch.
"""
import numpy as np
import pytest
import torch

from oracle import helmholtz as O

pytestmark = pytest.mark.gpu


def _cases(freqs, angles, interior_ref=96, pml=16, seed=3):
    from nsrno_amd.batch import PhysicalMedium, setup_case
    medium = PhysicalMedium(seed=seed, f_ref_mhz=4.0, interior_ref=interior_ref, pml=pml)
    out = []
    for f in freqs:
        case = medium.case(f)
        plan, ports, rhs, setup = setup_case(medium, case, angles, restart=3, autotune=False)
        out.append((case, plan, ports, rhs, setup))
    return out


def _close(built):
    for b in built:
        b[1].close()


def test_batch_equals_single_plan_solves():
    from nsrno_amd.helmholtz import solve_batch
    angles = np.linspace(-15.0, 15.0, 5)
    built = _cases([4.0, 4.9, 6.1], angles)
    try:
        sizes = [b[1].nz for b in built]
        print(f"grids {sizes}")
        assert len(set(sizes)) == 3
        xs, infos = solve_batch([b[1] for b in built], [b[3] for b in built], tolerance=1e-6, restart=3,
                                max_iterations=3000, graph=True)
        for (case, plan, ports, rhs, setup), xb, ib in zip(built, xs, infos):
            xs1, is1 = plan.solve(rhs, tolerance=1e-6, restart=3, max_iterations=3000, precond="mg", graph=True)
            d = float(torch.linalg.vector_norm(xb - xs1) / torch.linalg.vector_norm(xs1))
            print(f"{plan.nz}^2: steps batch {[i.iterations for i in ib]} single {[i.iterations for i in is1]}, "
                  f"field difference {d:.1e}")
            assert all(i.reason == "converged" for i in ib)
            assert [i.iterations for i in ib] == [i.iterations for i in is1]
            assert [i.applications for i in ib] == [i.applications for i in is1]
            assert [i.relative_residual for i in ib] == [i.relative_residual for i in is1]
            assert torch.equal(xb, xs1)
    finally:
        _close(built)


def test_batch_fields_match_sparse_lu():
    from nsrno_amd.helmholtz import solve_batch
    angles = np.linspace(-15.0, 15.0, 3)
    built = _cases([4.0, 5.3], angles)
    try:
        xs, infos = solve_batch([b[1] for b in built], [b[3] for b in built], tolerance=1e-8, restart=3,
                                max_iterations=3000)
        for (case, plan, ports, rhs, setup), x, info in zip(built, xs, infos):
            assert all(i.reason == "converged" and i.relative_residual < 1e-8 for i in info)
            work, weights, amp = setup["work"], setup["weights"], setup["amplitudes"]
            og = O.Grid(work.x_mm, work.z_mm, work.pml_mm)
            b = -(weights @ amp.T) / og.h_m ** 2
            ref, _ = O.solve(og, work.frequency_mhz, None, work.rho0 * work.c0 ** 2 / (work.rho * work.c ** 2) - 1,
                             work.rho0 / work.rho - 1, rhs=b, c0=work.c0, check_pml=False)
            got = x.cpu().numpy().reshape(len(angles), -1).T
            err = np.linalg.norm(got - ref, axis=0) / np.linalg.norm(ref, axis=0)
            print(f"{plan.nz}^2: field error vs sparse LU max {err.max():.2e}")
            assert err.max() < 1e-5
    finally:
        _close(built)


def test_batch_plans_leave_one_by_one():
    """A low and a high frequency need different numbers of cycles; the low
    one leaves the tables first.  Each plan's result equals its result when
    solved alone in a batch of one."""
    from nsrno_amd.helmholtz import solve_batch
    angles = np.linspace(-15.0, 15.0, 4)
    built = _cases([4.0, 7.5], angles, interior_ref=112)
    try:
        xs, infos = solve_batch([b[1] for b in built], [b[3] for b in built], tolerance=1e-7, restart=3)
        cyc = [max(i.cycles for i in info) for info in infos]
        print(f"cycles per plan {cyc}")
        assert cyc[0] != cyc[1]
        for k, b in enumerate(built):
            x1, i1 = solve_batch([b[1]], [b[3]], tolerance=1e-7, restart=3)
            assert [i.iterations for i in i1[0]] == [i.iterations for i in infos[k]]
            assert torch.equal(x1[0], xs[k])
    finally:
        _close(built)


def test_batched_sweep_config3_extremes():
    from nsrno_amd.batch import run_sweep
    freqs = np.linspace(4.0, 8.9, 64)[[0, 63]]
    angles = np.linspace(-15.0, 15.0, 11)
    cases_b, wall_b = run_sweep(freqs, angles, tol=1e-6, restart=3, workers=2, graph=True, mode="batched")
    assert [c.interior for c in cases_b] == [384, 854]
    for c in cases_b:
        assert c.reasons == ("converged",) and c.relative_residual < 1e-6
    cases_t, wall_t = run_sweep(freqs, angles, tol=1e-6, restart=3, workers=2, graph=True, autotune=False)
    print(f"batched {wall_b:.2f} s, threads {wall_t:.2f} s, steps "
          f"{[max(c.iterations) for c in cases_b]} / {[max(c.iterations) for c in cases_t]}")
    for cb, ct in zip(cases_b, cases_t):
        assert max(cb.iterations) == max(ct.iterations)
        np.testing.assert_array_equal(cb.traces, ct.traces)


def test_batch_argument_errors():
    from nsrno_amd._lib import HelmError
    from nsrno_amd.helmholtz import solve_batch
    angles = np.linspace(-15.0, 15.0, 3)
    built = _cases([4.0, 5.0], angles)
    try:
        p0, p1 = built[0][1], built[1][1]
        with pytest.raises(ValueError):
            solve_batch([p0, p1], [built[0][3]])
        with pytest.raises(ValueError):
            solve_batch([p0, p1], [built[0][3], built[1][3][:2]])
        with pytest.raises(HelmError, match="appears twice"):
            solve_batch([p0, p0], [built[0][3], built[0][3]])
    finally:
        _close(built)
