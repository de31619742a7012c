/* fc_lbm.h — C ABI of the MI355X-native Lattice Boltzmann engine (libfc_lbm.so).
 *
 * This is the drop-in boundary for FlowCraft's LbmSolver step
 * (FlowCraft/LbmSolver.h:59-302).  Every entry point is a plain C function over
 * host pointers and sizes; device memory is owned by the library.  All calls
 * return an int status (FC_LBM_OK = 0) unless stated otherwise; on failure
 * fc_lbm_last_error() describes the cause.  A handle is not thread-safe: the
 * reference serialises every solver call under one mutex (FlowCraft/main.cpp:
 * 366-373) and callers of this ABI must do the same.
 *
 * Field layout (FlowCraft/LbmSolver.h:190-214, consumers FieldOverlay.cpp:38-46
 * and SmokeTrail.cpp:96): 4 floats per cell at global index
 * cell = (z*ny + y)*nx + x; fluid cells hold (rho, ux, uy, uz), solid cells hold
 * (|F|/(6*sum w), Fx, Fy, Fz) — the momentum-exchange force of the CURRENT
 * populations (the LbmReference convention, LbmReference.cpp:127-148,475-500).
 */
#ifndef FLOWCRAFT_FC_LBM_H
#define FLOWCRAFT_FC_LBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FC_LBM_OK 0
#define FC_LBM_EINVAL 1    /* bad argument or size mismatch */
#define FC_LBM_ENOMEM 2    /* device allocation failed */
#define FC_LBM_EDEVICE 3   /* HIP runtime error */
#define FC_LBM_ESTATE 4    /* call not valid in the handle's state (e.g. no geometry yet) */

/* LatticeModel (FlowCraft/LatticeModel.h:12): same numbering as the reference enum. */
#define FC_LBM_D3Q15 0
#define FC_LBM_D3Q19 1

/* Population storage.  F32 mirrors LbmReference's float populations; F16/BF16 store
 * g = (f - w_i) * S (S = 4096 for F16, 1 for BF16) and run the collision in FP32;
 * F64 runs the whole step in double (the GPU twin of the FP64 oracle).  The _SR
 * variants store the same formats but round the step's stores stochastically (with
 * deterministic random bits per cell, step and direction) instead of to nearest, so
 * that changes smaller than the format's resolution survive on average; seeding,
 * restore and readback are as in the round-to-nearest format. */
#define FC_LBM_STORE_F32 0
#define FC_LBM_STORE_F16 1
#define FC_LBM_STORE_BF16 2
#define FC_LBM_STORE_F64 3
#define FC_LBM_STORE_F16_SR 4
#define FC_LBM_STORE_BF16_SR 5

/* Population layout: two SoA copies (pull, ping-pong) or one SoA copy updated in
 * place (esoteric pull). */
#define FC_LBM_LAYOUT_PINGPONG 0
#define FC_LBM_LAYOUT_INPLACE 1

typedef struct fc_lbm fc_lbm;

typedef struct fc_lbm_desc {
    int nx, ny, nz;     /* lattice size; nx >= 4 */
    int model;          /* FC_LBM_D3Q15 / FC_LBM_D3Q19 */
    int storage;        /* FC_LBM_STORE_* */
    int layout;         /* FC_LBM_LAYOUT_* */
    int periodic_y;     /* extension: wrap y instead of the free-slip/no-slip wall */
    int periodic_z;     /* extension: wrap z */
    void* stream;       /* hipStream_t to launch on; NULL = the library creates its own */
} fc_lbm_desc;

/* LbmSolver::Params (FlowCraft/LbmSolver.h:61-100); defaults tau 0.8, u 0.05,
 * reverse 0, free-slip y/z 1, Cs 0.12 (fc_lbm_params_default). */
typedef struct fc_lbm_params {
    float tau;
    float inlet_velocity;
    int reverse_flow;
    int free_slip_y;
    int free_slip_z;
    float smagorinsky_cs;
} fc_lbm_params;

/* LbmSolver::Stats (FlowCraft/LbmSolver.h:172-184) plus LbmReference::Stats'
 * totalMass (FlowCraft/LbmReference.h:112-119).  Sums accumulate in FP64. */
typedef struct fc_lbm_stats {
    double avg_density, min_density, max_density;
    double avg_speed, max_speed;
    double force_x, force_y, force_z;
    double max_stress;
    double total_mass;
    long long fluid_cells;
    int has_nan;
} fc_lbm_stats;

/* Replaces LbmSolver::Init's allocation half (FlowCraft/LbmSolver.cpp:205-366).
 * The device is the HIP current device at the call.  Lattice limit: at most
 * 2^32 - 2^20 cells nx*ny*nz (about 1620^3; one work-item per cell stays within one
 * 32-bit grid dimension), and three padded z layers of populations under 2 GiB; larger
 * lattices fail with FC_LBM_EINVAL before any device call.  Memory bounds it first for
 * most formats: per cell, 2*Q*s population bytes ping-pong (Q*s in place) plus 6 bytes
 * (solid mask, flags, boundary-record index), so one 288 GB MI355X holds about 1.8e9 cells of
 * FP32 D3Q19 ping-pong, 3.5e9 in place or in FP16 ping-pong, and the full 4.29e9 in
 * FP16 in place.  (FlowCraft's own ceiling is lower: its solid mask is one
 * 4-byte-per-cell buffer under 4 GiB, about 1.07e9 cells, FlowCraft/LbmSolver.h:22-58.) */
int fc_lbm_create(const fc_lbm_desc* desc, fc_lbm** out);

/* Replaces LbmSolver::Shutdown (FlowCraft/LbmSolver.cpp:368-387). */
int fc_lbm_destroy(fc_lbm* h);

/* Replaces LbmSolver::LastError (FlowCraft/LbmSolver.h:110).  Never NULL. */
const char* fc_lbm_last_error(const fc_lbm* h);

void fc_lbm_params_default(fc_lbm_params* p);

/* Replaces LbmSolver::Init's geometry upload (FlowCraft/LbmSolver.cpp:331-363):
 * solid is nx*ny*nz bytes (VoxelGrid::solid, 1 = solid), sdf is nx*ny*nz floats
 * (VoxelGrid::signedDistance) or NULL for plain bounce-back everywhere.  Builds the
 * per-cell flags, the Bouzidi link table q = clamp(phi_f/(phi_f - phi_s), 0, 1)
 * (FlowCraft/LbmReference.cpp:104-112) and the solid-surface list. Host pointers. */
int fc_lbm_set_geometry(fc_lbm* h, const uint8_t* solid, const float* sdf);

/* Extension: the same from device buffers (N mask bytes, N SDF floats or NULL), e.g. the
 * outputs of fc_voxelize_device.  The buffers must be complete when the call is made: work
 * that writes them on another stream must have been synchronised (the call orders itself
 * only after the handle's own stream).  The mask stays on the device; a host copy is made
 * only if a host-side use asks for it (fc_lbm_solid_mask, fc_lbm_aero_forces). */
int fc_lbm_set_geometry_device(fc_lbm* h, const uint8_t* solid_device, const float* sdf_device);

/* Replaces LbmSolver::SolidMask (FlowCraft/LbmSolver.h:216-218): the geometry's solid mask
 * (N bytes, 1 = solid) into a host array, copied from the library's own device copy (so a
 * caller's later reuse of the buffers given to fc_lbm_set_geometry[_device] does not
 * change it).  Blocking. */
int fc_lbm_solid_mask(fc_lbm* h, uint8_t* out);

/* Replaces LbmSolver::SetParams (FlowCraft/LbmSolver.h:135): state untouched. */
int fc_lbm_set_params(fc_lbm* h, const fc_lbm_params* p);
int fc_lbm_get_params(const fc_lbm* h, fc_lbm_params* p);

/* Replaces LbmSolver::Reset's seeding (FlowCraft/LbmSolver.cpp:389-427) with the
 * current params: quiescent (rho 1, u 0) or mid-speed (u = u_in/2, Bernoulli rho). */
int fc_lbm_reset_uniform(fc_lbm* h, int seed_mid);

/* Replaces LbmSolver::ResetFromField (FlowCraft/LbmSolver.cpp:429-472): per-cell
 * equilibrium of field4 (4 floats per cell, solid cells -> (1,0,0,0)).  Host pointer. */
int fc_lbm_reset_from_field(fc_lbm* h, const float* field4);
/* Same, from a device pointer (the learned warm start hands its prediction over in HBM). */
int fc_lbm_reset_from_field_device(fc_lbm* h, const float* field4_device);

/* Replaces LbmSolver::Step (FlowCraft/LbmSolver.cpp:556-562): n substeps, enqueued
 * on the handle's stream (asynchronous, like the reference's command recording). */
int fc_lbm_step(fc_lbm* h, int n);

/* Replaces LbmSolver::StepRange (FlowCraft/LbmSolver.cpp:543-554): one substep of
 * the cells with z in [z0, z0+zcount); call fc_lbm_advance once after the chunks
 * covering [0, nz). */
int fc_lbm_step_range(fc_lbm* h, int z0, int zcount);

/* Replaces LbmSolver::AdvanceSubstep (FlowCraft/LbmSolver.h:170). */
int fc_lbm_advance(fc_lbm* h);

/* Replaces LbmSolver::ReadbackFieldRange (FlowCraft/LbmSolver.cpp:584-651): writes
 * the layers [z0, z0+zcount) into out (host, sized for the WHOLE field, 4*N floats)
 * at their global offsets.  Blocking. */
int fc_lbm_readback_field_range(fc_lbm* h, int z0, int zcount, float* out);

/* Replaces LbmSolver::ReadbackField (FlowCraft/LbmSolver.cpp:722-737): whole field, host. */
int fc_lbm_readback_field(fc_lbm* h, float* out);

/* No counterpart in the reference: fc_lbm_readback_field_range for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, clamping and result, bit for bit; it calls no hipMalloc/hipFree.  The field
 * is staged through a fixed 32 MiB device array of the library (one per GPU, held until
 * the process ends) in chunks of whole z layers, so any range works as long as ONE layer
 * fits: nx*ny*16 bytes <= 32 MiB, else FC_LBM_EINVAL.  Every handle on a GPU shares that
 * GPU's array: calls that overlap on one GPU (handles on different threads or streams)
 * race and must be serialised by the caller.  Blocking. */
int fc_lbm_readback_field_range_noalloc(fc_lbm* h, int z0, int zcount, float* out);

/* The whole field into a device buffer of 4*N floats (no host round trip). */
int fc_lbm_field_device(fc_lbm* h, float* out_device);

/* Extension: the field at n listed cells (device array of global cell indices
 * (z*ny + y)*nx + x, 32-bit: cells below 2^31) into a device buffer of 4*n floats: fluid cells as
 * fc_lbm_field_device gives them, solid cells and indices outside [0, N) zero.  Enqueued
 * on the handle's stream (asynchronous); for consumers that sample a few cells (smoke
 * trails). */
int fc_lbm_field_at(fc_lbm* h, const int* cells_device, long long n, float* out_device);

/* Field snapshots for interactive front ends, in place of FlowCraft's paced, z-chunked
 * readback (ReadbackFieldPacedGPU, FlowCraft/main.cpp:2488-2555), which exists so that
 * a readback never holds the solver's queue for more than ~3 ms.  Here the copy never
 * holds it: fc_lbm_snapshot_begin enqueues the field of the current state into a device
 * buffer on the solver's stream (a fraction of a step) and its copy to pinned host
 * memory on a second stream, and returns at once; steps enqueued afterwards run while
 * the 4 floats per cell travel.  fc_lbm_snapshot_wait returns that host copy (valid
 * until the next begin or destroy); fc_lbm_snapshot_poll tells whether it has landed.
 * A begin waits for the previous snapshot's copy. */
int fc_lbm_snapshot_begin(fc_lbm* h);
int fc_lbm_snapshot_poll(fc_lbm* h, int* ready);
int fc_lbm_snapshot_wait(fc_lbm* h, const float** field);

/* Replaces LbmSolver::ReadbackStats (FlowCraft/LbmSolver.cpp:739-788) and
 * LbmReference::ComputeStats (FlowCraft/LbmReference.cpp:502-585), reduced on the
 * device in a fixed order (deterministic).  Blocking. */
int fc_lbm_readback_stats(fc_lbm* h, fc_lbm_stats* out);

/* Momentum-exchange force only (the Cd/Cl path): out[0..2] = sum of Fx, Fy, Fz over
 * all solid cells in FP64, out[3] = max stress.  Blocking. */
int fc_lbm_forces(fc_lbm* h, double out[4]);

/* Same sums left in device memory (4 doubles) without synchronising, for loops that
 * sample the force every few steps. */
int fc_lbm_forces_device(fc_lbm* h, double* out_device);

/* FieldOverlay::AerodynamicForces (FlowCraft/FieldOverlay.h:186-195): world-frame
 * vectors, grid-y along world X and grid-x along world Y as VoxelGrid::CellCenter
 * (FlowCraft/VoxelGrid.h:49-51). */
typedef struct fc_lbm_aero {
    float force_x, force_y, force_z;  /* sums over solid cells, FP64 then rounded */
    float min_stress, max_stress;     /* colour range: 0 .. max positive stress (1e-6 if none) */
    float center_of_pressure[3];      /* |Fz|-weighted, else |F|-weighted, else solid centroid */
    float lift_arrow_base[3];         /* (cop.x, cop.y, top of the solid above it) */
    int has_nan;                      /* a solid cell's field value was non-finite */
} fc_lbm_aero;

/* Replaces FieldOverlay::ComputeAerodynamicForces (FlowCraft/FieldOverlay.cpp:469-621)
 * applied to ReadbackField(): the same quantities, computed on the device from the
 * current populations in FP64 with a fixed reduction order, so no 4-float-per-cell
 * field crosses to the host.  cell_size and origin[3] are the VoxelGrid's world frame.
 * Blocking. */
int fc_lbm_aero_forces(fc_lbm* h, float cell_size, const float origin[3], fc_lbm_aero* out);

/* Replaces LbmSolver::ProbeSliceHealth (FlowCraft/LbmSolver.cpp:653-720): *diverged
 * = 1 if any cell of layer z (clamped) is non-finite, or a fluid cell has rho <= 0 or
 * |u|^2 > 4.  Blocking. */
int fc_lbm_probe_health(fc_lbm* h, int z, int* diverged);

/* LbmReference::HasDiverged (FlowCraft/LbmReference.h:129): sticky flag set by any
 * step since the last reset.  Blocking. */
int fc_lbm_has_diverged(fc_lbm* h, int* diverged);

/* Exact state dump/restore (extension; the reference only has the lossy
 * ResetFromField): f[d*N + cell] as float (F64 storage: as double), host pointers. */
int fc_lbm_get_populations(fc_lbm* h, void* out);
int fc_lbm_set_populations(fc_lbm* h, const void* in);

/* ---------------------------------------------------------------- voxelizer
 * FlowCraft's mesh voxelizer on the device.  Meshes are world-space vertex positions
 * (3 floats each) and triangle indices (3 per triangle, outward counter-clockwise
 * winding); grids use the VoxelGrid conventions (grid-x along world Y, grid-y along
 * world X, cell (x, y, z) at index (z*ny + y)*nx + x). */

/* Voxelizer::Params (FlowCraft/Voxelizer.h:13-30), same defaults via
 * fc_vox_params_default: 640 cells along the longest padded axis, padding (0.3, 1.0,
 * 0.5) of the mesh extent per world axis. */
typedef struct fc_vox_params {
    int target_resolution;
    float padding[3];
} fc_vox_params;

typedef struct fc_vox_grid {
    int nx, ny, nz;   /* 0, 0, 0 for an empty mesh (the reference's empty VoxelGrid) */
    float cell_size;
    float origin[3];  /* world position of cell (0,0,0)'s min corner */
} fc_vox_grid;

int fc_vox_params_default(fc_vox_params* p);

/* The grid Voxelizer::Voxelize (FlowCraft/Voxelizer.cpp:128-166) lays over a mesh:
 * dimensions, cell size and origin from the padded bounds, so the caller can size the
 * outputs of fc_voxelize.  Host arrays; an out-of-range index is FC_LBM_EINVAL. */
int fc_voxelize_frame(const float* positions, long long nverts, const uint32_t* indices, long long nindices,
                      const fc_vox_params* params, fc_vox_grid* out);

/* Replaces Voxelizer::Voxelize (FlowCraft/Voxelizer.cpp:128-432): the solid mask
 * (nx*ny*nz bytes, 1 = solid) and the signed distance (nx*ny*nz floats, + in fluid, -
 * in solid, +-inf far from the surface) of the frame above, into host arrays,
 * identical to the reference's.  Computed on the current device; blocking.  Nothing
 * is written for an empty mesh. */
int fc_voxelize(const float* positions, long long nverts, const uint32_t* indices, long long nindices,
                const fc_vox_params* params, uint8_t* solid, float* sdf);

/* Extension: the same into device buffers of `capacity` cells each (capacity bytes of mask,
 * capacity floats of distance), for fc_lbm_set_geometry_device: mesh to running solver
 * without the grid crossing PCIe.  FC_LBM_EINVAL, before any device work, when capacity
 * is below the frame's nx*ny*nz.  Blocking: the buffers are complete on return. */
int fc_voxelize_device(const float* positions, long long nverts, const uint32_t* indices, long long nindices,
                       const fc_vox_params* params, uint8_t* solid_device, float* sdf_device, long long capacity);

/* Message of the calling thread's last voxelizer error. */
const char* fc_voxelize_last_error(void);

/* ---------------------------------------------------------------- smoke trails
 * FlowCraft's SmokeTrailSystem (FlowCraft/SmokeTrail.h:10-184) with its particles and
 * trails on the device, advected through the solver's field without reading it back.
 * Grid conventions as VoxelGrid (world X = grid y, world Y = grid x).  confine_axis:
 * -1 none, 0/1/2 = the grid-x/y/z plane at confine_index (SmokeTrailSystem's
 * PlaneConfinement).  Random picks use the reference's generator and distributions
 * (std::mt19937 seeded 12345), in the reference's order. */
typedef struct fc_smoke fc_smoke;

/* A particle system over a grid: dimensions, cell size, origin and solid mask (host,
 * nx*ny*nz bytes, copied; grids below 2^31 cells, FC_LBM_EINVAL otherwise).  On a
 * device failure *out still receives the handle, for
 * fc_smoke_last_error; destroy it either way. */
int fc_smoke_create(int nx, int ny, int nz, float cell_size, const float origin[3], const uint8_t* solid,
                    fc_smoke** out);
int fc_smoke_destroy(fc_smoke* s);
const char* fc_smoke_last_error(const fc_smoke* s);

/* Replaces SmokeTrailSystem::Reset (FlowCraft/SmokeTrail.cpp:257-275): particle_count
 * particles scattered over the grid (or the plane), trails of at most
 * max(trail_length, 2) points. */
int fc_smoke_reset(fc_smoke* s, int reverse_flow, int particle_count, int trail_length, int confine_axis,
                   int confine_index);

/* Replaces SmokeTrailSystem::ScatterRandom (FlowCraft/SmokeTrail.cpp:202-216). */
int fc_smoke_scatter(fc_smoke* s, int confine_axis, int confine_index);

/* Replaces SmokeTrailSystem::ShiftPlane (FlowCraft/SmokeTrail.cpp:218-255). */
int fc_smoke_shift_plane(fc_smoke* s, int axis, int index);

/* Replaces SmokeTrailSystem::Step (FlowCraft/SmokeTrail.cpp:277-345) with the field of
 * the solver's current state (same lattice), computed on the device.  Blocking. */
int fc_smoke_step(fc_smoke* s, fc_lbm* h, int reverse_flow, float mm_per_sec_per_lattice_unit, float dt_seconds,
                  int suspend_respawn, int confine_axis, int confine_index);

/* Same with a caller's field: device pointer, 4 floats per cell (fc_lbm_field_device),
 * complete when the call is made (the work that wrote it already synchronised).
 * Blocking: the field may be reused when the call returns. */
int fc_smoke_step_field(fc_smoke* s, const float* field_device, int reverse_flow, float mm_per_sec_per_lattice_unit,
                        float dt_seconds, int suspend_respawn, int confine_axis, int confine_index);

/* The trails, host: out[(p*trail_length + k)*3 + c] for k < lengths[p], oldest first. */
int fc_smoke_trails(fc_smoke* s, float* out, int* lengths);
int fc_smoke_info(const fc_smoke* s, int* particles, int* trail_length);

/* Waits for all work queued on the handle's stream. */
int fc_lbm_sync(fc_lbm* h);

/* Introspection used by bindings, tests and the bench. */
int fc_lbm_info(const fc_lbm* h, int* nx, int* ny, int* nz, int* q, int* storage, int* layout);
long long fc_lbm_step_count(const fc_lbm* h);
size_t fc_lbm_population_bytes(const fc_lbm* h);
void* fc_lbm_stream(const fc_lbm* h);

/* Boundary link table dump (tests pin q-fractions bit-exactly): up to max links,
 * each (cell, direction i, kind, q) for a fluid cell whose direction-i source is
 * solid.  kind: 2 plain bounce-back, 3 Bouzidi q < 1/2, 4 Bouzidi q >= 1/2.
 * *count receives the total number of links. */
int fc_lbm_boundary_links(fc_lbm* h, long long max, long long* cell, int* dir, int* kind, float* q,
                          long long* count);

/* Number of solid cells with a fluid neighbour (the force kernel's work list). */
long long fc_lbm_surface_cells(const fc_lbm* h);

#ifdef __cplusplus
}
#endif

#endif /* FLOWCRAFT_FC_LBM_H */
