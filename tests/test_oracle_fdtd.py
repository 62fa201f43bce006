"""Pins the FDTD oracle (oracle/fdtd.c) before anything is compared against it.

The reference ships no FDTD golden vectors (its inputs are time-seeded, graphfdtd.cpp:113), so
the oracle is pinned by checks that a plausible mistake -- a wrong sign, a swapped operand, an
off-by-one index or the wrong coefficient -- cannot pass:
  * an independent array-slicing formulation transcribed from the MATLAB lines the reference
    keeps beside each loop (graphfdtd.cpp:183,203,242,266,290) agrees bit for bit;
  * the Yee scheme's discrete Gauss laws: div H (every cell) and div E (interior nodes) are
    invariants of the exact recurrence, so in binary64 they drift only by rounding;
  * constant and zero fields are fixed points; the update is linear (exact for a scale of 2);
  * the tangential E faces are copied through unchanged (the reference's "PEC");
  * the coefficients follow run_graph_fdtd_task (graphfdtd.cpp:1357-1381).
"""
import numpy as np
import pytest

import oracle
from oracle import FIELDS


def shaped(a, f, nx, ny, nz):
    d1, d2, d3 = oracle.extent(f, nx, ny, nz)
    return a.reshape(d3, d2, d1)  # [k, j, i], i fastest


def numpy_step(F, p, dtype):
    """One step written with array slices ([k,j,i] indexing), from the reference's MATLAB forms."""
    t = dtype
    nx, ny, nz = p.Nx, p.Ny, p.Nz
    Hx, Hy, Hz, Ex, Ey, Ez = (shaped(F[f].astype(t), f, nx, ny, nz) for f in range(6))
    cH = t(p.Dt) / t(p.mu0)
    cE = t(p.Dt) / t(p.eps0)
    Cx, Cy, Cz = t(p.Cx), t(p.Cy), t(p.Cz)
    # Hx = Hx + (Dt/mu0)*((Ey(:,:,2:Nz+1)-Ey(:,:,1:Nz))*Cz - (Ez(:,2:Ny+1,:)-Ez(:,1:Ny,:))*Cy)
    Hx2 = Hx + cH * ((Ey[1:, :, :] - Ey[:-1, :, :]) * Cz - (Ez[:, 1:, :] - Ez[:, :-1, :]) * Cy)
    # Hy = Hy + (Dt/mu0)*((Ez(2:Nx+1,:,:)-Ez(1:Nx,:,:))*Cx - (Ex(:,:,2:Nz+1)-Ex(:,:,1:Nz))*Cz)
    Hy2 = Hy + cH * ((Ez[:, :, 1:] - Ez[:, :, :-1]) * Cx - (Ex[1:, :, :] - Ex[:-1, :, :]) * Cz)
    # Hz = Hz + (Dt/mu0)*((Ex(:,2:Ny+1,:)-Ex(:,1:Ny,:))*Cy - (Ey(2:Nx+1,:,:)-Ey(1:Nx,:,:))*Cx)
    Hz2 = Hz + cH * ((Ex[:, 1:, :] - Ex[:, :-1, :]) * Cy - (Ey[:, :, 1:] - Ey[:, :, :-1]) * Cx)
    # Ex(:,2:Ny,2:Nz) += (Dt/eps0)*((Hz(:,2:Ny,2:Nz)-Hz(:,1:Ny-1,2:Nz))*Cy
    #                               - (Hy(:,2:Ny,2:Nz)-Hy(:,2:Ny,1:Nz-1))*Cz)
    Ex2 = Ex.copy()
    Ex2[1:nz, 1:ny, :] = Ex[1:nz, 1:ny, :] + cE * (
        (Hz2[1:nz, 1:ny, :] - Hz2[1:nz, 0:ny - 1, :]) * Cy
        - (Hy2[1:nz, 1:ny, :] - Hy2[0:nz - 1, 1:ny, :]) * Cz)
    # Ey(2:Nx,:,2:Nz) += (Dt/eps0)*((Hx(2:Nx,:,2:Nz)-Hx(2:Nx,:,1:Nz-1))*Cz
    #                               - (Hz(2:Nx,:,2:Nz)-Hz(1:Nx-1,:,2:Nz))*Cx)
    Ey2 = Ey.copy()
    Ey2[1:nz, :, 1:nx] = Ey[1:nz, :, 1:nx] + cE * (
        (Hx2[1:nz, :, 1:nx] - Hx2[0:nz - 1, :, 1:nx]) * Cz
        - (Hz2[1:nz, :, 1:nx] - Hz2[1:nz, :, 0:nx - 1]) * Cx)
    # Ez(2:Nx,2:Ny,:) += (Dt/eps0)*((Hy(2:Nx,2:Ny,:)-Hy(1:Nx-1,2:Ny,:))*Cx
    #                               - (Hx(2:Nx,2:Ny,:)-Hx(2:Nx,1:Ny-1,:))*Cy)
    Ez2 = Ez.copy()
    Ez2[:, 1:ny, 1:nx] = Ez[:, 1:ny, 1:nx] + cE * (
        (Hy2[:, 1:ny, 1:nx] - Hy2[:, 1:ny, 0:nx - 1]) * Cx
        - (Hx2[:, 1:ny, 1:nx] - Hx2[:, 0:ny - 1, 1:nx]) * Cy)
    return [a.reshape(-1) for a in (Hx2, Hy2, Hz2, Ex2, Ey2, Ez2)]


SHAPES = [(4, 4, 4), (7, 5, 3), (1, 6, 2), (9, 1, 4), (5, 6, 1), (12, 10, 8)]


@pytest.mark.parametrize("shape", SHAPES)
def test_oracle_matches_slicing_formulation_bitwise(shape):
    nx, ny, nz = shape
    p = oracle.make_params(nx, ny, nz, 0.5)
    F = oracle.init_fields(nx, ny, nz, seed=1234 + nx)
    G = list(F)
    for _ in range(3):
        G = numpy_step(G, p, np.float32)
    R = oracle.fdtd_run(F, p, 3)
    for f in range(6):
        assert np.array_equal(R[f].view(np.uint32), G[f].view(np.uint32)), FIELDS[f]


def div_h(F, p):
    nx, ny, nz = p.Nx, p.Ny, p.Nz
    Hx, Hy, Hz = (shaped(F[f], f, nx, ny, nz) for f in range(3))
    return ((Hx[:, :, 1:] - Hx[:, :, :-1]) * p.Cx + (Hy[:, 1:, :] - Hy[:, :-1, :]) * p.Cy
            + (Hz[1:, :, :] - Hz[:-1, :, :]) * p.Cz)


def div_e_interior(F, p):
    nx, ny, nz = p.Nx, p.Ny, p.Nz
    Ex, Ey, Ez = (shaped(F[f], f, nx, ny, nz) for f in range(3, 6))
    return ((Ex[1:nz, 1:ny, 1:nx] - Ex[1:nz, 1:ny, 0:nx - 1]) * p.Cx
            + (Ey[1:nz, 1:ny, 1:nx] - Ey[1:nz, 0:ny - 1, 1:nx]) * p.Cy
            + (Ez[1:nz, 1:ny, 1:nx] - Ez[0:nz - 1, 1:ny, 1:nx]) * p.Cz)


@pytest.mark.parametrize("shape", [(6, 5, 4), (9, 7, 8)])
def test_discrete_gauss_laws_hold_in_binary64(shape):
    nx, ny, nz = shape
    p = oracle.make_params(nx, ny, nz, 0.5)
    F = [a.astype(np.float64) for a in oracle.init_fields(nx, ny, nz, seed=77)]
    R = oracle.fdtd_run(F, p, 10, dtype=np.float64)
    dh0, dh1 = div_h(F, p), div_h(R, p)
    de0, de1 = div_e_interior(F, p), div_e_interior(R, p)
    # Scale: the size of one curl term; the divergence itself is O(C) for random fields.
    sh = np.abs(dh0).max()
    se = np.abs(de0).max()
    assert np.abs(dh1 - dh0).max() <= 1e-12 * sh
    assert np.abs(de1 - de0).max() <= 1e-12 * max(se, np.abs(R[3]).max() * p.Cx)
    # and the fields did change a lot, so the invariant is not vacuous
    assert np.abs(R[3] - F[3]).max() > 1.0


def test_constant_and_zero_fields_are_fixed_points():
    nx, ny, nz = 5, 4, 3
    p = oracle.make_params(nx, ny, nz, 0.5)
    F = [np.full(np.prod(oracle.extent(f, nx, ny, nz)), 0.375 * (f + 1), np.float32)
         for f in range(6)]
    R = oracle.fdtd_run(F, p, 4)
    for f in range(6):
        assert np.array_equal(R[f], F[f])
    Z = [np.zeros_like(a) for a in F]
    R = oracle.fdtd_run(Z, p, 4)
    assert all(not a.any() for a in R)


def test_linearity_exact_for_power_of_two_scale():
    nx, ny, nz = 6, 5, 4
    p = oracle.make_params(nx, ny, nz, 0.5)
    F = oracle.init_fields(nx, ny, nz, seed=5)
    R1 = oracle.fdtd_run(F, p, 3)
    R2 = oracle.fdtd_run([np.float32(2) * a for a in F], p, 3)
    for f in range(6):
        assert np.array_equal(np.float32(2) * R1[f], R2[f])


def test_tangential_e_faces_copied_through():
    nx, ny, nz = 6, 5, 4
    p = oracle.make_params(nx, ny, nz, 0.5)
    F = oracle.init_fields(nx, ny, nz, seed=9)
    R = oracle.fdtd_run(F, p, 5)
    Ex0, Ex1 = shaped(F[3], 3, nx, ny, nz), shaped(R[3], 3, nx, ny, nz)
    for sl in [(0, slice(None), slice(None)), (nz, slice(None), slice(None)),
               (slice(None), 0, slice(None)), (slice(None), ny, slice(None))]:
        assert np.array_equal(Ex0[sl], Ex1[sl])
    Ey0, Ey1 = shaped(F[4], 4, nx, ny, nz), shaped(R[4], 4, nx, ny, nz)
    for sl in [(0, slice(None), slice(None)), (nz, slice(None), slice(None)),
               (slice(None), slice(None), 0), (slice(None), slice(None), nx)]:
        assert np.array_equal(Ey0[sl], Ey1[sl])
    Ez0, Ez1 = shaped(F[5], 5, nx, ny, nz), shaped(R[5], 5, nx, ny, nz)
    for sl in [(slice(None), 0, slice(None)), (slice(None), ny, slice(None)),
               (slice(None), slice(None), 0), (slice(None), slice(None), nx)]:
        assert np.array_equal(Ez0[sl], Ez1[sl])
    # interior E did change
    assert not np.array_equal(Ex0[1:nz, 1:ny, :], Ex1[1:nz, 1:ny, :])


def test_single_update_argument_order():
    """Each half-step reads (own, in1, in2) in the reference's port order (graphfdtd.cpp:448-560):
    Hx(Hx,Ey,Ez) Hy(Hy,Ex,Ez) Hz(Hz,Ex,Ey) Ex(Ex,Hy,Hz) Ey(Ey,Hx,Hz) Ez(Ez,Hx,Hy)."""
    nx, ny, nz = 5, 4, 3
    p = oracle.make_params(nx, ny, nz, 0.5)
    F = oracle.init_fields(nx, ny, nz, seed=11)
    G = numpy_step(F, p, np.float32)
    H = [oracle.fdtd_update(0, F[0], F[4], F[5], p), oracle.fdtd_update(1, F[1], F[3], F[5], p),
         oracle.fdtd_update(2, F[2], F[3], F[4], p)]
    E = [oracle.fdtd_update(3, F[3], H[1], H[2], p), oracle.fdtd_update(4, F[4], H[0], H[2], p),
         oracle.fdtd_update(5, F[5], H[0], H[1], p)]
    for f, a in enumerate(H + E):
        assert np.array_equal(a, G[f]), FIELDS[f]


def test_coefficients_follow_reference_and_courant():
    p = oracle.make_params(64, 64, 64, 0.01)
    f = np.float32
    # reference: nrm = sqrt(Cx^2+Cy^2+Cz^2)*100, Dt = 1/(c0*nrm)  (graphfdtd.cpp:1361-1380)
    cx, cy, cz = f(64) / f(0.05), f(64) / f(0.04), f(64) / f(0.03)
    assert (p.Cx, p.Cy, p.Cz) == (float(cx), float(cy), float(cz))
    nrm_ref = np.sqrt(f(cx * cx + cy * cy + cz * cz), dtype=f) * f(100)
    assert abs(p.Dt - float(f(1) / (f(299792458.0) * nrm_ref))) <= 1e-6 * p.Dt
    q = oracle.make_params(64, 64, 64, 0.5)
    # Courant number c0*Dt*sqrt(Cx^2+Cy^2+Cz^2) is what the workload asks for
    s = q.c0 * q.Dt * np.sqrt(q.Cx ** 2 + q.Cy ** 2 + q.Cz ** 2)
    assert abs(s - 0.5) < 1e-6
    assert abs(q.mu0 - 4e-7 * 3.1415) < 1e-12 and q.eps0 == float(f(8.8541878e-12))


def test_fp32_vs_fp64_drift_is_small_at_courant_half():
    nx, ny, nz = 16, 12, 10
    p = oracle.make_params(nx, ny, nz, 0.5)
    F = oracle.init_fields(nx, ny, nz, seed=3)
    R32 = oracle.fdtd_run(F, p, 20)
    R64 = oracle.fdtd_run(F, p, 20, dtype=np.float64)
    for f in range(6):
        err = np.abs(R32[f] - R64[f]).max() / np.abs(R64[f]).max()
        assert err < 1e-4, (FIELDS[f], err)
