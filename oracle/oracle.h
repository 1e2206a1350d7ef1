/*
 * Oracle: a CPU FP64 restatement of Mara3's Godunov block update.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the parity reference
 * (or as the timed CPU baseline).  The product path (mara3p_amd) never calls
 * it and fails loudly when its HIP library is missing.
 *
 * Every function restates one reference function, cited file:line, in the
 * reference's own association order (no FMA contraction: build with
 * -ffp-contract=off).  Arrays use the reference's logical layout: row-major,
 * last axis fastest, one tuple of n_c doubles per zone in logical component
 * order (Euler D,S1,S2,S3,E / rho,v1,v2,v3,p; iso2d Sigma,Sx,Sy / Sigma,vx,vy).
 *
 * Pinning (how each function is checked; see tests/test_oracle_*.py):
 *   plm, euler/iso2d/mhd/srhd pointwise physics, two_body : reference headers
 *       compiled into oracle/_ref (bitwise) + reference KATs.
 *   update2d skeleton + minidisk update : reference minidisk_scheme.cpp
 *       compiled into oracle/_ref (golden fixtures, <=1e-12).
 *   euler1d/euler2d module update : shares the update2d/update1d skeleton
 *       pinned above, plus uniform-state, conservation and exact-Sod checks.
 *   guard fill (extend) : brute-force index identity on tiny meshes.
 *   mhd CT step : reference scheme_mhd_v1.cpp compiled into oracle/_ref.
 */
#ifndef MARA3P_ORACLE_H
#define MARA3P_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- pointwise ---------------- */
double orc_plm(double yl, double y0, double yr, double theta);

int  orc_euler_recover_primitive(const double* u, double gamma, double* p);
void orc_euler_conserved_density(const double* p, double gamma, double* u);
void orc_euler_flux(const double* p, int axis, double gamma, double* f);
void orc_euler_riemann_hlle(const double* pl, const double* pr, int axis, double gamma, double* f);
void orc_euler_riemann_hllc(const double* pl, const double* pr, int axis, double gamma, double* f);

void orc_iso2d_recover_primitive(const double* u, double* p);
void orc_iso2d_conserved_density(const double* p, double* u);
void orc_iso2d_riemann_hlle(const double* pl, const double* pr, double cs2, int axis, double* f);

/* two-body: state = {M1,x1,y1,vx1,vy1, M2,x2,y2,vx2,vy2} after rotation by
 * -omega_frame*t (problems/minidisk_scheme.cpp:333-335). Returns Newton iters. */
int  orc_two_body_state(double a, double M, double q, double e, double t, double omega_frame, double* state);
void orc_two_body_acceleration(double mass, double px, double py, double x, double y, double rs, double* a);

/* ---------------- minidisk solver data (problems/minidisk.hpp:80-96) ---------------- */
typedef struct {
    int    block_size;
    int    depth;
    double omega_frame;
    double mach_number;
    double kinematic_viscosity;
    double domain_radius;
    double softening_length;
    double eccentricity;
    double mass_ratio;
    double sink_radius;
    double sink_rate;
    double buffer_rate;
    double buffer_scale;
    double cfl;
} orc_minidisk_t;

/* ---------------- block updates (module level) ---------------- */
int orc_euler1d_updated_conserved(const double* uc, const double* pe, int n,
                                  double dt, double dl, double theta, double gamma,
                                  int outflow_lower, int outflow_upper, double* out);
int orc_euler2d_updated_conserved(const double* uc, const double* pe, int ni, int nj,
                                  double dt, double dl, double theta, double gamma, int hllc, double* out);
int orc_locally_isothermal_updated_conserved(const double* uc, const double* pe, int ni, int nj,
                                  double dt, double dl, const double* cs2_x, const double* cs2_y, double nu, double* out);
int orc_minidisk_updated_conserved(const double* uc, const double* pe, const orc_minidisk_t* sd,
                                  int level, long bi, long bj, double time, double dt, double* out);
void orc_minidisk_initial_conserved(const orc_minidisk_t* sd, int level, long bi, long bj, double* out);
void orc_minidisk_cell_coordinates(const orc_minidisk_t* sd, int level, long bi, long bj, double* xc, double* yc);

/* ---------------- guard fill on a uniform periodic block mesh ---------------- */
/* P: all blocks' arrays [nb*nb][n][n][nc] (block b = bi*nb+bj); pe: [n+4][n+4][nc] */
void orc_extend_block_2d(const double* P, int nb, int n, int nc, long bi, long bj, double* pe);
/* P: [nb][n][nc]; pe: [n+4][nc] */
void orc_extend_block_1d(const double* P, int nb, int n, int nc, long bi, double* pe);

/* ---------------- mesh-level RK2 steps (in place), OpenMP over blocks ---------------- */
int orc_euler1d_mesh_rk2(double* u, int nb, int n, double dt, double dl, double theta, double gamma,
                         int nsteps, int nthreads);
int orc_euler2d_mesh_rk2(double* u, int nb, int n, double dt, double dl, double theta, double gamma,
                         int hllc, int nsteps, int nthreads);
int orc_minidisk_mesh_rk2(double* u, const orc_minidisk_t* sd, double t0, double dt, int nsteps, int nthreads);
/* one explicit-Euler stage over the mesh (the reference minidisk driver's substep) */
int orc_minidisk_mesh_stage(const double* u, double* out, const orc_minidisk_t* sd, double time, double dt, int nthreads);
int orc_euler2d_mesh_stage(const double* u, double* out, int nb, int n, double dt, double dl, double theta,
                           double gamma, int hllc, int nthreads);

/* ---------------- ideal MHD with constrained transport (oracle_mhd.c) ---------------- */
/* primitive (rho, v1, v2, v3, p, B1, B2, B3); conserved (D, S1, S2, S3, E) */
int    orc_mhd_recover_primitive(const double* u, const double* b, double gamma, double* p);
void   orc_mhd_conserved_density(const double* p, double gamma, double* u);
void   orc_mhd_flux(const double* p, int axis, double gamma, double* f);
double orc_mhd_fastest_speed(const double* p, double gamma);
void   orc_mhd_riemann_hlle(const double* pl, const double* pr, double bl, int axis, double gamma, double* f, double* e);
/* advance_mhd_ct on a periodic N^3 cube (uc [N][N][N][5], bf1 [N+1][N][N],
 * bf2 [N][N+1][N], bf3 [N][N][N+1]); dt from the CFL condition */
int orc_mhd_advance_ct(int N, const double* uc, const double* bf1, const double* bf2, const double* bf3,
                       double dl, double cfl, double theta, double gamma,
                       double* uo, double* bo1, double* bo2, double* bo3, double* dt_out);
int orc_mhd_rk2(int N, double* uc, double* bf1, double* bf2, double* bf3, double dl, double cfl, double theta,
                double gamma, double* dts);
/* mhd_scheme_v2 array functions (src/scheme_mhd_v2.cpp:75-278), adiabatic index 5/3 */
int  orc_mhd_v2_primitive_array(int n1, int n2, int n3, const double* uc, const double* bf1, const double* bf2,
                                const double* bf3, double* pc);
void orc_mhd_v2_godunov_fluxes(int m1, int m2, int m3, const double* pc, const double* bf1, const double* bf2,
                               const double* bf3, double* gf1, double* gf2, double* gf3);
void orc_mhd_v2_electromotive_forces(int m1, int m2, int m3, const double* gf1, const double* gf2, const double* gf3,
                                     double* ee1, double* ee2, double* ee3);
void orc_mhd_v2_updated_conserved_density(int n1, int n2, int n3, const double* uc, const double* gf1,
                                          const double* gf2, const double* gf3, double dt, double dl, double* out);
void orc_mhd_v2_curl(int n1, int n2, int n3, const double* e1, const double* e2, const double* e3, double dl,
                     double* b1, double* b2, double* b3);
void orc_mhd_v2_updated_magnetic_flux_density(int n1, int n2, int n3, const double* bf1, const double* bf2,
                                              const double* bf3, const double* ee1, const double* ee2,
                                              const double* ee3, double dt, double dl, double* b1, double* b2,
                                              double* b3);
void orc_mhd_v2_extend_cells(const double* const* blocks, int bs, int count, int nq, double* out);
void orc_mhd_v2_extend_faces(const double* const* blocks, int bs, int count, double* bf1, double* bf2, double* bf3);

/* ---------------- special-relativistic hydrodynamics (oracle_srhd.c) ---------------- */
/* primitive (rho, u1, u2, u3, p), u = gamma beta; conserved (D, S1, S2, S3, tau) */
int  orc_srhd_recover_primitive(const double* u, double gamma, double temperature_floor, double* p);
void orc_srhd_conserved_density(const double* p, double gamma, double* u);
void orc_srhd_flux(const double* p, int axis, double gamma, double* f);
/* mode 0 HLLE, 1 HLLC, 2 contact speed (f[0]), 3 HLLC through a face moving at face_speed */
int  orc_srhd_riemann(const double* pl, const double* pr, int axis, double gamma, int mode, double face_speed, double* f);

/* exact Riemann solution of the 1D Euler equations (Toro ch. 4) for checks */
void orc_euler_exact_riemann(const double* pl, const double* pr, double gamma, double xi, double* p);

#ifdef __cplusplus
}
#endif
#endif
