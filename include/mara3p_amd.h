/*
 * mara3p_amd: C ABI of the MI355X Godunov block update.
 *
 * All array arguments are DEVICE pointers (HIP, gfx950) unless marked
 * "host".  Layout is structure-of-arrays per block, row-major with the last
 * axis fastest, exactly the reference ndarray block with the per-zone tuple
 * split into components in logical order:
 *     state   [nblk][nc][n][n]          (2D)      [nblk][nc][n]   (1D)
 *     pe      [nc][n+4][n+4]            primitive block with 2 guard zones
 * Euler nc = 5 (D, S1, S2, S3, E / rho, v1, v2, v3, p); iso2d nc = 3.
 *
 * Every entry point returns an m3p status code.  Stream arguments are
 * hipStream_t passed as void*.  Stage entry points are asynchronous on that
 * stream; errors raised inside their kernels (e.g. negative density) are
 * latched in a slot of that stream and reported by m3p_check_errors(stream).
 * The per-block entry points are blocking: each call latches its kernels'
 * errors in a slot of its own and returns the status before returning (so
 * they behave like the reference's throwing functions).
 *
 * Threads: any number of host threads may call the library at once.  The
 * per-block calls are safe on any stream, the same one included (the
 * reference calls updated_conserved concurrently, one block per thread-pool
 * worker, src/parallel_computable.cpp:248-296).  The asynchronous calls keep
 * their error latch and reduction scratch per stream, so concurrent callers
 * use one stream each; calls on one stream are ordered by the stream.  A
 * stream's state is created by its first call, which must come before any
 * graph capture on it (m3p_graph_begin makes it).  Graph replays report
 * kernel errors to the latch of the stream they were captured on.  The tile
 * table (m3p_set_stage2d_tile, m3p_autotune_*) is read and written
 * atomically.
 */
#ifndef MARA3P_AMD_H
#define MARA3P_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define M3P_OK                     0
#define M3P_ERR_NEGATIVE_DENSITY   1   /* euler::recover_primitive, src/physics_euler.hpp:158-161 */
#define M3P_ERR_NEGATIVE_PRESSURE  2   /* mhd::recover_primitive, src/physics_mhd.hpp:296-300 */
#define M3P_ERR_NO_CONVERGENCE     3   /* srhd::recover_primitive, src/physics_srhd.hpp:273-277 */
#define M3P_ERR_BAD_SHAPE          4   /* nd::select / shape errors, src/core_ndarray.hpp:714-732 */
#define M3P_ERR_HIP                5
#define M3P_ERR_BAD_ARGUMENT       6

#define M3P_KIND_EULER2D   0   /* stage2d scheme kinds, for tile selection */
#define M3P_KIND_MINIDISK  1
#define M3P_KIND_AMR_EULER2D 2

/* Precision modes.  M3P_FP64 is the drop-in: FP64 in the reference's
 * association order.  Its parity bar per scheme: Euler (1D, 2D, quad-tree,
 * HLLE and HLLC), iso2d / locally isothermal, SRHD and the mhd_scheme_v2
 * array functions equal the reference bit for bit; minidisk and MHD
 * constrained transport agree within 1e-12 relative (minidisk: the device
 * exp, reciprocal forms of its sources and Newton-refined hardware
 * reciprocals and square roots differ from glibc and from the reference's
 * divisions by a few ulp; MHD: the same forms in the wave speeds, the HLLE
 * flux and the primitive recovery).  M3P_MIXED computes primitives, slopes
 * and fluxes in FP32 and keeps the conserved state and flux differences in
 * FP64. */
#define M3P_FP64   0
#define M3P_MIXED  1

/* Library / error handling ------------------------------------------------ */
int         m3p_version(void);
const char* m3p_last_error(void);       /* text of the last failure on this thread */
/* Synchronise `stream`, read and clear its error latch.  *code is the kernel
 * status (M3P_OK if none), block/zone locate the first offender. */
int         m3p_check_errors(void* stream, int* code, int* block, int* zone);

/* Per-block entry points (compatibility path) ---------------------------- */

/* modules::euler2d::updated_conserved, src/module_euler.cpp:263-298
 * (declared src/module_euler.hpp:182-190).  uc [5][n][n], pe [5][n+4][n+4]
 * primitives, out [5][n][n].  hllc = 0 reproduces the reference (HLLE). */
int m3p_euler2d_updated_conserved(const double* uc, const double* pe, double* out, int n,
                                  double dt, double dl, double plm_theta, double gamma_law_index,
                                  int hllc, int precision, void* stream);

/* modules::euler1d::updated_conserved, src/module_euler.cpp:107-135, with the
 * outflow boundary of src/mara.cpp:123-143 selected by the two flags (the
 * reference applies only the lower one when a block is both first and last). */
int m3p_euler1d_updated_conserved(const double* uc, const double* pe, double* out, int n,
                                  double dt, double dl, double plm_theta, double gamma_law_index,
                                  int outflow_lower, int outflow_upper, int precision, void* stream);

/* Minidisk solver data, problems/minidisk.hpp:80-96. */
typedef struct {
    int    block_size;
    int    depth;
    double omega_frame;
    double mach_number;
    double kinematic_viscosity;   /* viscous stress when != 0 (per-zone kernel) */
    double domain_radius;
    double softening_length;
    double eccentricity;
    double mass_ratio;
    double sink_radius;
    double sink_rate;
    double buffer_rate;
    double buffer_scale;
    double cfl;
} m3p_minidisk_data;

/* minidisk::updated_conserved, problems/minidisk_scheme.cpp:294-355
 * (declared problems/minidisk.hpp:164-170).  uc [3][n][n] conserved
 * densities, pe [3][n+4][n+4] primitives, block (level, bi, bj). */
int m3p_minidisk_updated_conserved(const double* uc, const double* pe, double* out,
                                   int level, int bi, int bj, double time, double dt,
                                   const m3p_minidisk_data* data /* host */, int precision, void* stream);

/* locally_isothermal::updated_conserved, src/module_locally_isothermal.cpp:182-212,
 * with the viscous stress of godunov_and_viscous_fluxes (:106-174) when
 * kinematic_viscosity != 0 (then pe's corner zones are read too).  The
 * reference's temperature mapping is a host callback; callers evaluate it at
 * mesh_geometry.face_coordinates and pass cs2_x [(n+1)][n] and cs2_y
 * [n][(n+1)] (device). */
int m3p_locally_isothermal_updated_conserved(const double* uc, const double* pe, double* out, int n,
                                   double dt, double dl, const double* cs2_x, const double* cs2_y,
                                   double kinematic_viscosity, int precision, void* stream);

/* locally_isothermal::updated_conserved over every block of a uniform mesh
 * (the per-block update of updated_conserved_tree,
 * src/module_locally_isothermal.cpp:218-238, with its guard fill fused as in
 * m3p_euler2d_stage): nc = 3, nbr as m3p_euler2d_stage, face sound speeds
 * squared per block cs2_x [nblk][(n+1)][n], cs2_y [nblk][n][(n+1)] (device),
 * viscous stress when kinematic_viscosity != 0.  Asynchronous; the RK
 * combine and dt_dev as m3p_euler2d_stage. */
int m3p_locally_isothermal_stage(const double* u, const double* u0, double* out, const int* nbr,
                                 int nblk, int n, double dt, const double* dt_dev, double dl, double rk_b,
                                 const double* cs2_x, const double* cs2_y, double kinematic_viscosity,
                                 int precision, void* stream);

/* Block helpers of the module API --------------------------------------- */
/* recover_primitive_array, src/module_euler.cpp:84-89 (euler1d) and :240-245
 * (euler2d): u, p [5][nzones]; blocking; negative density reported like the
 * reference's throw (src/physics_euler.hpp:158-161). */
int m3p_euler_recover_primitive_array(const double* u, double* p, long nzones, double gamma_law_index, void* stream);
/* recover_primitive_array, src/module_locally_isothermal.cpp:86-91: [3][nzones]. */
int m3p_iso2d_recover_primitive_array(const double* u, double* p, long nzones, void* stream);
/* estimate_gradient, src/module_euler.cpp:251-257: PLM slopes of p
 * [nc][ni][nj] along `axis` (0 or 1) into g [nc][ni-2][nj] or [nc][ni][nj-2]. */
int m3p_plm_gradient_2d(const double* p, double* g, int nc, int ni, int nj, int axis, double plm_theta, void* stream);
/* weighted_sum / weighted_sum_tree, src/mesh_amr.hpp:455-467: out = s*b + t*(1-b). */
int m3p_weighted_sum(const double* s, const double* t, double* out, long n, double b, void* stream);

/* Whole-mesh stage entry points (throughput path) ------------------------- */
/* One RK stage over every block of a mesh, guard fill fused: zones outside a
 * block are read from its neighbours through nbr [nblk][4] (device) =
 * {prev axis 0, next axis 0, prev axis 1, next axis 1} (the periodic
 * next_on/prev_on of src/mesh_block_index.hpp:334-335 for uniform meshes).
 * out = u0 * rk_b + L(u) * (1 - rk_b) when u0 != NULL, else L(u)
 * (amr::weighted_sum_tree, src/mesh_amr.hpp:455-467; RK2 uses rk_b = 0.5).
 * dt_dev, when non-NULL, overrides dt with a device-resident value. */
int m3p_euler2d_stage(const double* u, const double* u0, double* out, const int* nbr,
                      int nblk, int n, double dt, const double* dt_dev, double dl, double rk_b,
                      double plm_theta, double gamma_law_index, int hllc, int precision, void* stream);

/* Minidisk stage: block coordinates bcoord [nblk][2] (device), all blocks at
 * data->depth.  The binary's state at `time` is solved on the host
 * (two_body::orbital_state + rotate, src/physics_two_body.hpp:109-157,417-421). */
int m3p_minidisk_stage(const double* u, const double* u0, double* out, const int* nbr, const int* bcoord,
                       int nblk, double time, double dt, double rk_b,
                       const m3p_minidisk_data* data /* host */, int precision, void* stream);

/* A minidisk step with its clock in device memory, so that whole steps can be
 * captured in a hipGraph and replayed (the stage above takes the time from
 * the host).  m3p_minidisk_clock: the time step rule of the reference's
 * minidisk driver (problems/minidisk.cpp:202-206: smallest_cell_size / vmax
 * * cfl * min(1, 0.01 + t / 0.05)) and the binary's state
 * (src/physics_two_body.hpp:109-157,417-421) evaluated by one device thread.
 * clock [2] (device doubles) = {t, dt}: reads t, writes dt into clock[1],
 * the binary's state {M1, x1, y1, M2, x2, y2} at t into orbit[0..5] and at
 * t + dt into orbit[6..11] (device doubles), and the time after the step
 * into clock[0]: t * 0.5 + (t + dt) * 0.5 for an RK2 step (rk_order 2, the
 * solution-time combination of src/module_euler.cpp:357), t + dt for the
 * reference driver's forward-Euler substep (rk_order 1,
 * problems/minidisk.cpp:172-186).  The device's sin and cos may differ from
 * the host's in the last bit, so the orbit may differ from the host stage's
 * by an ulp.  m3p_minidisk_stage_dev: m3p_minidisk_stage with the binary's
 * state read from orbit (6 device doubles: orbit for the first stage of a
 * step, orbit + 6 for the second) and dt from dt_dev (clock + 1). */
int m3p_minidisk_clock(const m3p_minidisk_data* data /* host */, double* clock, double* orbit, int rk_order,
                       void* stream);
int m3p_minidisk_stage_dev(const double* u, const double* u0, double* out, const int* nbr, const int* bcoord,
                           int nblk, const double* orbit, const double* dt_dev, double rk_b,
                           const m3p_minidisk_data* data /* host */, int precision, void* stream);

/* One whole RK2 step of the minidisk mesh in one launch (no reference entry
 * point: the step of control::advance_runge_kutta, src/app_control.hpp:
 * 169-195, over minidisk::updated_conserved): out = u * rk_b +
 * L(L(u)) * (1 - rk_b) with the stages at time and time + dt, equal value
 * for value to m3p_minidisk_stage(u -> u1, time) followed by
 * m3p_minidisk_stage(u1, u -> out, time + dt), without storing u1.  Uniform
 * meshes of blocks of 64, 128 or 256 zones a side, M3P_MIXED only, no
 * viscosity (otherwise M3P_ERR_BAD_ARGUMENT or M3P_ERR_BAD_SHAPE); out must
 * not alias u.  The _dev form reads the binary's state of both stages from
 * orbit [12] and dt from dt_dev, as m3p_minidisk_clock writes them. */
int m3p_minidisk_rk2_step(const double* u, double* out, const int* nbr, const int* bcoord, int nblk,
                          double time, double dt, double rk_b, const m3p_minidisk_data* data /* host */,
                          int precision, void* stream);
int m3p_minidisk_rk2_step_dev(const double* u, double* out, const int* nbr, const int* bcoord, int nblk,
                              const double* orbit, const double* dt_dev, double rk_b,
                              const m3p_minidisk_data* data /* host */, int precision, void* stream);

/* 1D Euler stage over nblk blocks of n zones, out-flow at the two ends of the
 * chain described by nbr [nblk][2] = {prev, next} (-1 = outflow boundary). */
int m3p_euler1d_stage(const double* u, const double* u0, double* out, const int* nbr,
                      int nblk, int n, double dt, double dl, double rk_b,
                      double plm_theta, double gamma_law_index, int precision, void* stream);

/* Ideal MHD with constrained transport ------------------------------------ */
/* mara::advance_mhd_ct, src/scheme_mhd_v1.cpp:265-299 (declared
 * src/scheme_mhd_v1.hpp:108-118), periodic cube of N^3 cells, PLM + HLLE,
 * edge EMFs averaged from face EMFs, dt = cfl * dl / max fastest speed
 * computed by this call (written to dt, device double, may be NULL).
 *   uc, uo [5][N][N][N]; bf1, bo1 [N+1][N][N]; bf2, bo2 [N][N+1][N];
 *   bf3, bo3 [N][N][N+1]   (the reference's face arrays, both periodic end
 *   faces stored).
 * With u0 != NULL the RK combine out = base * rk_b + advance * (1 - rk_b) is
 * fused (examples/mhd3d.cpp:77-86 weighted_sum).  work: device scratch of at
 * least m3p_mhd_workspace_bytes(N, precision) bytes (primitives, face fluxes
 * and face fields).  Asynchronous; negative density / pressure are latched
 * (M3P_ERR_NEGATIVE_DENSITY / _PRESSURE, src/physics_mhd.hpp:296-300). */
long m3p_mhd_workspace_bytes(int N, int precision);
int  m3p_mhd_advance_ct(const double* uc, const double* bf1, const double* bf2, const double* bf3,
                        const double* u0, const double* b01, const double* b02, const double* b03, double rk_b,
                        int N, double dl, double cfl_number, double plm_theta, double gamma_law_index,
                        double* uo, double* bo1, double* bo2, double* bo3, double* dt,
                        void* work, long work_bytes, int precision, void* stream);
/* m3p_mhd_advance_ct for callers that cannot allocate (no reference
 * counterpart: for code under hipGraph capture, tight serving loops and
 * harnesses that forbid hipMalloc).  The same kernels and results as
 * m3p_mhd_advance_ct, with no workspace argument: the workspace is a static
 * device array of 48 MiB inside the library (one per GPU, held until the
 * process ends), so N is limited to what fits (m3p_mhd_workspace_bytes(N,
 * precision) <= 48 MiB: N a multiple of 8, whose passes keep no face fluxes,
 * up to 72 in FP64 and 96 mixed; any other N up to 59 and 75); a larger N fails with
 * M3P_ERR_BAD_ARGUMENT.  Kernel errors (negative density or pressure) go to
 * status, a caller's device int[4] (code, block, zone) that the call clears
 * first; m3p_check_errors does not see them.  Nothing is allocated, with
 * hipMalloc, torch or otherwise.  All calls on one GPU share the one array,
 * which keeps the last call's contents: calls that overlap on one GPU (other
 * streams or threads) race and must be serialised by the caller. */
int  m3p_mhd_advance_ct_noalloc(const double* uc, const double* bf1, const double* bf2, const double* bf3,
                                const double* u0, const double* b01, const double* b02, const double* b03, double rk_b,
                                int N, double dl, double cfl_number, double plm_theta, double gamma_law_index,
                                double* uo, double* bo1, double* bo2, double* bo3, double* dt, int* status,
                                int precision, void* stream);
/* max over cells of |div B| = |d1 B1 + d2 B2 + d3 B3| (divergence_difference,
 * src/mesh_cartesian_3d.hpp:290-301) into out (device double). */
int  m3p_mhd_max_div_b(const double* bf1, const double* bf2, const double* bf3, int N, double* out, void* stream);

/* mhd_scheme_v2 array functions (src/scheme_mhd_v2.hpp:71-210, bodies
 * src/scheme_mhd_v2.cpp:75-278) over one block, FP64 in the reference's
 * arithmetic, adiabatic index 5/3 (src/scheme_mhd_v2.cpp:48).  Device arrays,
 * row-major, components outermost: cell tuples [nq][s1][s2][s3] (5 conserved,
 * 8 primitive (rho, v1, v2, v3, p, B1, B2, B3), 8 Godunov data = 5 fluxes +
 * 3 face electric field components); face and edge arrays are scalars in the
 * reference's shapes: for cells (n1, n2, n3), x faces (n1+1, n2, n3), x edges
 * (n1, n2+1, n3+1), and the other axes by symmetry.  Asynchronous on stream,
 * except m3p_mhd_v2_primitive_array, which waits and returns the status. */

/* primitive_array :75-81 (face_to_cell, src/mesh_cartesian_3d.hpp:235-245):
 * uc [5] (n1, n2, n3) + faces -> pc [8] (n1, n2, n3).  Blocking;
 * M3P_ERR_NEGATIVE_DENSITY / _PRESSURE where mhd::recover_primitive throws
 * (src/physics_mhd.hpp:296-300), for one offending cell. */
int m3p_mhd_v2_primitive_array(const double* uc, const double* bf1, const double* bf2, const double* bf3,
                               int n1, int n2, int n3, double* pc, void* stream);
/* godunov_fluxes :168-185 (first-order face states, HLLE): pc [8] (m1, m2, m3)
 * extended by one on every axis; bf1 (m1-1, m2, m3), bf2 (m1, m2-1, m3),
 * bf3 (m1, m2, m3-1) -> gf1, gf2, gf3 [8] of the same face shapes. */
int m3p_mhd_v2_godunov_fluxes(const double* pc, const double* bf1, const double* bf2, const double* bf3,
                              int m1, int m2, int m3, double* gf1, double* gf2, double* gf3, void* stream);
/* electromotive_forces :191-214: gf of cells (m1, m2, m3) -> edge fields
 * ee1 (m1, m2-1, m3-1), ee2 (m1-1, m2, m3-1), ee3 (m1-1, m2-1, m3). */
int m3p_mhd_v2_electromotive_forces(const double* gf1, const double* gf2, const double* gf3, int m1, int m2, int m3,
                                    double* ee1, double* ee2, double* ee3, void* stream);
/* updated_conserved_density :220-241: uc [5] (n1, n2, n3); gf extended by one
 * on the transverse axes (gf1 (n1+1, n2+2, n3+2), ...) -> out [5] (n1, n2, n3). */
int m3p_mhd_v2_updated_conserved_density(const double* uc, const double* gf1, const double* gf2, const double* gf3,
                                         int n1, int n2, int n3, double dt, double dl, double* out, void* stream);
/* curl :247-258: edges e1 (n1, n2+1, n3+1), e2, e3 -> faces over a unit time. */
int m3p_mhd_v2_curl(const double* ee1, const double* ee2, const double* ee3, int n1, int n2, int n3, double dl,
                    double* bf1, double* bf2, double* bf3, void* stream);
/* updated_magnetic_flux_density :264-278: faces of cells (n1, n2, n3); edges as
 * electromotive_forces returns them for gf extended by one (ee1 (n1+2, n2+1,
 * n3+1), ...; the longitudinal end layers are not read) -> bo1, bo2, bo3. */
int m3p_mhd_v2_updated_magnetic_flux_density(const double* bf1, const double* bf2, const double* bf3,
                                             const double* ee1, const double* ee2, const double* ee3,
                                             int n1, int n2, int n3, double dt, double dl,
                                             double* bo1, double* bo2, double* bo3, void* stream);
/* extend_cell_primitive_variables :126-135: pc_blocks is a host array of 27
 * device pointers, the blocks [8] (bs, bs, bs) of a 3 x 3 x 3 neighbourhood in
 * row-major order of the offsets (mesh::neighbors_27,
 * src/mesh_cartesian_3d.hpp:355-364) -> out [8] (bs+2c, bs+2c, bs+2c),
 * 0 <= count <= bs. */
int m3p_mhd_v2_extend_cell_primitive_variables(const double* const* pc_blocks, int block_size, int count,
                                               double* out, void* stream);
/* extend_face_magnetic_flux_density :141-162: bf_blocks is a host array of 27
 * device pointers; entries 9a .. 9a+8 are the axis-a face arrays of the 3 x 3
 * square of blocks normal to axis a (mesh::neighbors_9, :384-395, row-major)
 * -- the component the reference takes from each of its 27 face triplets ->
 * bf1 (bs+1, bs+2c, bs+2c), bf2 (bs+2c, bs+1, bs+2c), bf3 (bs+2c, bs+2c, bs+1). */
int m3p_mhd_v2_extend_face_magnetic_flux_density(const double* const* bf_blocks, int block_size, int count,
                                                 double* bf1, double* bf2, double* bf3, void* stream);

/* Special-relativistic hydrodynamics (pointwise, FP64) -------------------- */
/* srhd::recover_primitive, src/physics_srhd.hpp:224-279 (Newton on the
 * pressure, <= 50 iterations, tolerance 1e-12): u, p [5][nzones]; blocking;
 * M3P_ERR_NO_CONVERGENCE / M3P_ERR_NEGATIVE_PRESSURE as the reference throws. */
int m3p_srhd_recover_primitive_array(const double* u, double* p, long nzones, double gamma_law_index,
                                     double temperature_floor, void* stream);
/* srhd::conserved_density, :285-299: p, u [5][nzones]. */
int m3p_srhd_conserved_density_array(const double* p, double* u, long nzones, double gamma_law_index, void* stream);
/* srhd::riemann_solver (Mignone & Bodo 2005), :377-496, per face:
 * pl, pr, f [5][nfaces]; mode 0 = riemann_hlle, 1 = riemann_hllc (face at
 * rest), 2 = contact speed (written to f[0][i]), 3 = riemann_hllc through a
 * face moving at face_speed.  Blocking. */
int m3p_srhd_riemann_array(const double* pl, const double* pr, double* f, long nfaces, int axis,
                           double gamma_law_index, int mode, double face_speed, void* stream);

/* Tile selection and autotuning (no reference counterpart: the reference runs
 * one closure per block on a CPU thread pool).  A 2D stage workgroup owns TJ
 * contiguous columns of a block and marches down H rows; the compiled
 * variants are TJ x H in {256x32, 256x16, 128x32, 64x64, 64x32, 64x16, 32x32,
 * 16x16}.  The choice is per process, per (kind, precision). */
int m3p_set_stage2d_tile(int kind, int precision, int tj, int h);
int m3p_get_stage2d_tile(int kind, int precision, int* tj /* host */, int* h /* host */);
/* Time every variant that fits on the caller's mesh (u -> out, `reps`
 * launches each, blocking), keep the fastest and report it (host outputs). */
int m3p_autotune_euler2d(const double* u, double* out, const int* nbr, int nblk, int n, double dl,
                         double gamma_law_index, int precision, int reps, int* tj, int* h, float* ms, void* stream);
int m3p_autotune_minidisk(const double* u, double* out, const int* nbr, const int* bcoord, int nblk,
                          const m3p_minidisk_data* data /* host */, int precision, int reps,
                          int* tj, int* h, float* ms, void* stream);

/* Batched and quad-tree meshes -------------------------------------------- */
/* 1D Euler stage over a work-list of blocks of n zones that may belong to
 * many independent problems (an uneven ensemble): chain nbr [nblk][2] as
 * m3p_euler1d_stage (-1 = outflow end), cell spacing per block dl_blk
 * [nblk] (device).  One persistent launch walks the list; dt_dev (device,
 * may be NULL) overrides dt. */
int m3p_euler1d_stage_blocks(const double* u, const double* u0, double* out, const int* nbr, const double* dl_blk,
                             int nblk, int n, double dt, const double* dt_dev, double rk_b,
                             double plm_theta, double gamma_law_index, int precision, void* stream);
/* euler2d stage over the leaves of a 2:1-balanced quad-tree, with the guard
 * fill across level jumps of amr::extend_block over amr::get_or_create_block
 * (src/mesh_amr.hpp:300-393: same level copy, coarse neighbour injection,
 * fine neighbours 2x2 averaged, all on primitives) and cell spacing dl_leaf
 * [nleaf] per leaf (cell_spacing, src/mesh_geometry.cpp:168-171).
 * pe_work NULL: one fused launch reads guard zones through the recipes.
 * pe_work [nleaf][5][n+4][n+4]: the extended primitive blocks are written
 * there first (the reference's extend-then-update structure), then the
 * update reads them.  Both give identical results.  n even, >= 16.  recipes
 * [nleaf][4][8] int (directions -x, +x, -y, +y):
 * {kind 0 same / 1 coarse / 2 fine, leaf, child00, child10, child01,
 *  child11, orthant_i, orthant_j}.  As the reference, no flux correction at
 * level jumps. */
int m3p_amr_euler2d_stage(const double* u, const double* u0, double* out, double* pe_work, const int* recipes,
                          const double* dl_leaf, int nleaf, int n, double dt, const double* dt_dev, double rk_b,
                          double plm_theta, double gamma_law_index, int hllc, int precision, void* stream);
/* The quad-tree stage above in a conservative mode (no reference
 * counterpart: the reference applies no flux correction at level jumps,
 * src/mesh_amr.hpp:300-328, problems/minidisk.cpp:60-62).  Fused guard fill
 * as m3p_amr_euler2d_stage with pe_work NULL; the stage also writes the face
 * fluxes of every leaf edge to flux_work [nleaf][4][5][n] (device, caller's;
 * edges -x, +x, -y, +y), then every zone next to an edge bordering finer
 * leaves is corrected so that its flux through that edge is the mean of the
 * two finer leaves' face fluxes covering each face (the mass, momentum and
 * energy leaving one side enter the other, to round-off).  Zones away from
 * such edges are unchanged. */
int m3p_amr_euler2d_stage_conservative(const double* u, const double* u0, double* out, double* flux_work,
                                       const int* recipes, const double* dl_leaf, int nleaf, int n, double dt,
                                       const double* dt_dev, double rk_b, double plm_theta, double gamma_law_index,
                                       int hllc, int precision, void* stream);
/* CFL with a cell spacing per block: dt = cfl / max over zones of
 * (max(|v1|, |v2|) + c_s) / dl_b (the reference fixes dt, src/mara.cpp:305-308;
 * this time step is this library's). */
int m3p_euler_cfl_dt_blocks(const double* u, int nblk, long zones_per_block, const double* dl_blk,
                            double gamma_law_index, double cfl, double* dt_out, void* stream);

/* One whole RK2 step of a uniform euler2d mesh in one launch (no reference
 * entry point: the step of euler2d::updated_solution,
 * src/module_euler.cpp:304-323, without storing its intermediate state):
 * out = u * rk_b + L(L(u)) * (1 - rk_b), equal value for value to
 * m3p_euler2d_stage(u -> u1) followed by m3p_euler2d_stage(u1, u -> out)
 * with the same dt (or dt_dev), dl, theta, gamma and hllc.  Blocks of 64,
 * 128 or 256 zones a side, M3P_MIXED only (otherwise M3P_ERR_BAD_ARGUMENT
 * or M3P_ERR_BAD_SHAPE); out must not alias u.  The _cfl form also writes
 * the time step of `out` to dt_next (device double), as
 * m3p_euler2d_stage_cfl does for its output, bit for bit. */
int m3p_euler2d_rk2_step(const double* u, double* out, const int* nbr, int nblk, int n, double dt,
                         const double* dt_dev, double dl, double rk_b, double plm_theta, double gamma_law_index,
                         int hllc, int precision, void* stream);
int m3p_euler2d_rk2_step_cfl(const double* u, double* out, const int* nbr, int nblk, int n, double dt,
                             const double* dt_dev, double dl, double rk_b, double plm_theta, double gamma_law_index,
                             int hllc, double cfl, double* dt_next, int precision, void* stream);

/* One whole RK2 step of the 1D work-list in one launch (no reference entry
 * point: two euler1d::updated_conserved passes, src/module_euler.cpp:107-135,
 * and the RK combine, without storing the intermediate state): equal value
 * for value to m3p_euler1d_stage_blocks(u -> u1) followed by
 * m3p_euler1d_stage_blocks(u1, u -> out) with the same dt (or dt_dev),
 * dl_blk, theta and gamma.  The prev/next links of nbr must be mutual where
 * both blocks exist (chains and rings; otherwise the error latch receives
 * M3P_ERR_BAD_SHAPE); blocks of 16, 32, 64, 128 or a multiple of 256 zones,
 * M3P_MIXED only (otherwise M3P_ERR_BAD_ARGUMENT or M3P_ERR_BAD_SHAPE); out
 * must not alias u.  The _cfl form also writes the time step of `out` to
 * dt_next (device double), as m3p_euler1d_stage_blocks_cfl does. */
int m3p_euler1d_rk2_step_blocks(const double* u, double* out, const int* nbr, const double* dl_blk, int nblk, int n,
                                double dt, const double* dt_dev, double rk_b, double plm_theta,
                                double gamma_law_index, int precision, void* stream);
int m3p_euler1d_rk2_step_blocks_cfl(const double* u, double* out, const int* nbr, const double* dl_blk, int nblk,
                                    int n, double dt, const double* dt_dev, double rk_b, double plm_theta,
                                    double gamma_law_index, double cfl, double* dt_next, int precision, void* stream);

/* The three Euler stages above with the CFL reduction of their output fused
 * in (no reference counterpart: a device step driver's CFL step without a
 * separate pass over the state).  Same arguments and results as
 * m3p_euler2d_stage / m3p_euler1d_stage_blocks / m3p_amr_euler2d_stage,
 * and dt_next (device double) receives the time step m3p_euler_cfl_dt /
 * m3p_euler_cfl_dt_blocks would compute from `out` with this cfl, bit for
 * bit.  dt_next is written after the stage has read dt_dev, so the two may be
 * the same buffer.  The reduction slots are those of `stream`. */
int m3p_euler2d_stage_cfl(const double* u, const double* u0, double* out, const int* nbr,
                          int nblk, int n, double dt, const double* dt_dev, double dl, double rk_b,
                          double plm_theta, double gamma_law_index, int hllc, double cfl, double* dt_next,
                          int precision, void* stream);
int m3p_euler1d_stage_blocks_cfl(const double* u, const double* u0, double* out, const int* nbr, const double* dl_blk,
                                 int nblk, int n, double dt, const double* dt_dev, double rk_b,
                                 double plm_theta, double gamma_law_index, double cfl, double* dt_next,
                                 int precision, void* stream);
int m3p_amr_euler2d_stage_cfl(const double* u, const double* u0, double* out, double* pe_work, const int* recipes,
                              const double* dl_leaf, int nleaf, int n, double dt, const double* dt_dev, double rk_b,
                              double plm_theta, double gamma_law_index, int hllc, double cfl, double* dt_next,
                              int precision, void* stream);

/* hipGraph capture (no reference counterpart): wrap any sequence of
 * asynchronous m3p calls issued on `stream` between m3p_graph_begin and
 * m3p_graph_end; replay with m3p_graph_launch.  Time steps that change
 * between replays must come from device memory (the dt_dev arguments, e.g.
 * written by m3p_euler_cfl_dt inside the captured sequence). */
int m3p_graph_begin(void* stream);
int m3p_graph_end(void* stream, void** graph_exec /* host out */);
int m3p_graph_launch(void* graph_exec, void* stream);
int m3p_graph_destroy(void* graph_exec);

/* Reductions --------------------------------------------------------------- */
/* CFL time step for Euler states: dt = cfl / max over zones of
 * (max(|v1|, |v2|) + c_s) / dl, written to dt_out (device double).  The
 * reduction is a wave64 shuffle max per workgroup and an atomic max across
 * workgroups (positive doubles order like their bit patterns), so it is
 * deterministic. */
int m3p_euler_cfl_dt(const double* u, int nblk, long zones_per_block, double gamma_law_index,
                     double cfl, double dl, double* dt_out, void* stream);

/* FP64 totals of each conserved component over the mesh:
 * totals[q] = sum over blocks and zones of u[b][q][...] (device double[nc]). */
int m3p_conserved_totals(const double* u, int nblk, int nc, long zones_per_block, double* totals, void* stream);

#ifdef __cplusplus
}
#endif
#endif
