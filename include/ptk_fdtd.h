/*
 * ptk_fdtd.h -- the FDTD constant block passed by value to every Yee update task, and the
 * staggered-grid extents of the six field arrays.
 *
 * FDTD_PARAMS keeps the reference layout field for field (ptaskUnitTest/graphfdtd.h:6-19):
 * graphs push it through a PTPARM_BYVALSTRUCT template (graphfdtd.cpp:88) into the sticky
 * "Config" port at formal index 3 (graphfdtd.cpp:421), so its 40 bytes are part of the kernel ABI.
 *
 * Arrays are i-fastest: v(i,j,k) = a[i + j*d1 + k*d1*d2] (ptaskUnitTest/Simple3DArray.h:28-35),
 * with extents (graphfdtd.cpp:70-86):
 *   Hx (Nx+1, Ny,   Nz  )   Hy (Nx,   Ny+1, Nz  )   Hz (Nx,   Ny,   Nz+1)
 *   Ex (Nx,   Ny+1, Nz+1)   Ey (Nx+1, Ny,   Nz+1)   Ez (Nx+1, Ny+1, Nz  )
 */
#ifndef PTK_FDTD_H
#define PTK_FDTD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct _fdtdcb_t {
    int Nx;
    int Ny;
    int Nz;

    float Cx;
    float Cy;
    float Cz;

    float eps0; /* permittivity of vacuum */
    float mu0;  /* permeability of vacuum */
    float c0;   /* speed of light in vacuum */
    float Dt;   /* time step */
} FDTD_PARAMS;

/* Field identifiers, in the order the six update tasks of one step run. */
enum { PTK_FDTD_HX = 0, PTK_FDTD_HY = 1, PTK_FDTD_HZ = 2,
       PTK_FDTD_EX = 3, PTK_FDTD_EY = 4, PTK_FDTD_EZ = 5 };

#if defined(__HIPCC__)
#define PTK_FDTD_FN static inline __host__ __device__
#else
#define PTK_FDTD_FN static inline
#endif

/* Extents (d1,d2,d3) of field f on an Nx x Ny x Nz grid. */
PTK_FDTD_FN void ptk_fdtd_extent(int f, int Nx, int Ny, int Nz, int *d1, int *d2, int *d3) {
    int x = Nx, y = Ny, z = Nz;
    switch (f) {
        case PTK_FDTD_HX: x = Nx + 1; break;
        case PTK_FDTD_HY: y = Ny + 1; break;
        case PTK_FDTD_HZ: z = Nz + 1; break;
        case PTK_FDTD_EX: y = Ny + 1; z = Nz + 1; break;
        case PTK_FDTD_EY: x = Nx + 1; z = Nz + 1; break;
        case PTK_FDTD_EZ: x = Nx + 1; y = Ny + 1; break;
        default: break;
    }
    *d1 = x; *d2 = y; *d3 = z;
}

PTK_FDTD_FN unsigned long long ptk_fdtd_field_elems(int f, int Nx, int Ny, int Nz) {
    int a, b, c;
    ptk_fdtd_extent(f, Nx, Ny, Nz, &a, &b, &c);
    return (unsigned long long)a * (unsigned long long)b * (unsigned long long)c;
}

/*
 * Coefficients for an Nx x Ny x Nz cavity of 0.05 x 0.04 x 0.03 m, following
 * run_graph_fdtd_task (ptaskUnitTest/graphfdtd.cpp:1357-1381): Cx = Nx/Lx etc., eps0 =
 * 8.8541878e-12, mu0 = 4e-7 * 3.1415 (pi truncated as in the reference), c0 = 299792458.
 * The reference fixes Dt = 1/(c0*nrm) with nrm = 100*sqrt(Cx^2+Cy^2+Cz^2) (Courant number
 * 0.01); `courant` generalises that: Dt = courant/(c0*sqrt(Cx^2+Cy^2+Cz^2)), so courant = 0.01
 * reproduces the reference and courant = 0.5 is the measured workload.  All arithmetic is in
 * float, as in the reference, so host and device see the same 40 bytes.
 */
static inline FDTD_PARAMS ptk_fdtd_make_params(int Nx, int Ny, int Nz, float courant) {
    FDTD_PARAMS p;
    const float Lx = 0.05f, Ly = 0.04f, Lz = 0.03f;
    p.Nx = Nx;
    p.Ny = Ny;
    p.Nz = Nz;
    p.Cx = (float)Nx / Lx;
    p.Cy = (float)Ny / Ly;
    p.Cz = (float)Nz / Lz;
    p.eps0 = 8.8541878e-12f;
    p.mu0 = 4e-7f * 3.1415f;
    p.c0 = 299792458.0f;
    {
        float s = p.Cx * p.Cx + p.Cy * p.Cy + p.Cz * p.Cz;
        /* sqrtf is correctly rounded (IEEE 754), so this is portable bit for bit */
        float nrm = __builtin_sqrtf(s);
        p.Dt = courant / (p.c0 * nrm);
    }
    return p;
}

#ifdef __cplusplus
}
#endif

#endif /* PTK_FDTD_H */
