"""The adjoint-gradient workload: the 64x64x16 two-material geometry with
per-voxel copper conductivity uniform in [3e7, 6e7] S/m from the seed,
J = sum over its 32 frequencies (1e7..1e10 Hz) of |Z_11|^2.

The full dJ/dsigma by the adjoint solve (M^-H K^H GMRES) is held to
  - the reciprocity shortcut (no adjoint solve) and the gradient built on
    the plain FP64 path's own forward solutions (the reference's solve);
  - central differences with relative step 1e-4 at 16 seeded conductor
    voxels, each J from a batched FP64 sweep at tol 1e-12.
The central difference resolves J to about 1e-11 absolute through the
solve tolerance, so agreement is stated relative to the largest gradient
entry; entries above 1e-7 also agree to 1e-2 relative."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(900)
def test_adjoint_gradient_workload():
    from solver_amd import gradcheck, peec
    with peec.Context(0) as ctx:
        r = gradcheck.run(nfreq=32, nvox=16, tol=1e-12, step=1e-4, fd_path="batched", seed=0, ctx=ctx)
    print("J %.10e (plain %.10e); adjoint %.2f s; vs reciprocity %.2e, vs plain-path gradient %.2e; "
          "central differences: max |fd - g| / max |g| = %.2e" % (r["J"], r["J_plain"], r["adjoint_gradient_s"],
                                                                   r["adjoint_vs_reciprocity"], r["adjoint_vs_plain"],
                                                                   r["fd_max_err_of_max"]), flush=True)
    for row in r["fd"]:
        print("  voxel %6d: adjoint %+.6e  FD %+.6e  rel %.2e" % (row["voxel"], row["adjoint"], row["fd"],
                                                                  row["err_rel"]), flush=True)
    assert r["adjoint_flags"] == [0]
    assert r["frequencies"] == 32 and len(r["fd"]) == 16
    assert abs(r["J"] - r["J_plain"]) <= 1e-9 * abs(r["J"])
    assert r["adjoint_vs_reciprocity"] <= 1e-9
    assert r["adjoint_vs_plain"] <= 1e-9
    assert r["fd_max_err_of_max"] <= 1e-5
    for row in r["fd"]:
        if abs(row["adjoint"]) > 1e-7:
            assert row["err_rel"] <= 1e-2
