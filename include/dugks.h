/*
 * dugks.h -- C-ABI of the MI355X (gfx950) DUGKS time-step engine.
 *
 * The engine advances the Shakhov-model discrete unified gas kinetic scheme
 * (DUGKS) on an unstructured finite-volume mesh.  It replaces the body of
 * Foam::fvDVM (src/fvDVM/fvDVM/fvDVM.H:78-377 in dugksFoam): an OpenFOAM-side
 * shim keeps the fvDVM.H public API and forwards to the functions below.
 *
 * Conventions
 *  - Plain pointers and sizes only; no C++ types cross the ABI, no exceptions.
 *  - Every function returns an int status: 0 = ok, < 0 = dugks_status code.
 *    dugks_last_error(h) returns a message for the last failure on h.
 *  - Caller-owned host buffers are copied during the call and never retained.
 *  - One engine = one device + one HIP stream.  Calls on one engine must be
 *    serialised by the caller; get_* calls synchronise the engine's stream.
 *  - Mesh numbering follows OpenFOAM's polyMesh: faces 0..n_internal_faces-1
 *    are internal (owner < neighbour, Sf points owner -> neighbour), then the
 *    faces of every non-empty boundary patch in patch order (Sf outward).
 *    Empty patches (2-D front/back) are omitted, as fvPatch gives them size 0.
 *  - Discrete velocities are the tensor product of the 1-D set in
 *    constant/Xis + constant/weights, ordered k = (iz*nY + iy)*nX + ix with
 *    xi = (Xis[ix], Xis[iy], Xis[iz]) and w = w[iz]*w[iy]*w[ix]
 *    (fvDVM.C:140-221; z components are 0 in 2-D, y and z components in 1-D,
 *    where nv = nDV and w = w[ix]).
 */
#ifndef DUGKS_H
#define DUGKS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DUGKS_ABI_VERSION 2

/* status codes */
enum dugks_status {
    DUGKS_OK = 0,
    DUGKS_ERR_INVALID = -1,     /* invalid argument or case description   */
    DUGKS_ERR_DEVICE = -2,      /* HIP runtime error                       */
    DUGKS_ERR_MEMORY = -3,      /* device allocation failed                */
    DUGKS_ERR_STATE = -4,       /* call not valid in the engine's state    */
    DUGKS_ERR_NONFINITE = -5,   /* step produced a non-finite or rho <= 0  */
    DUGKS_ERR_UNSUPPORTED = -6  /* feature not supported by this build     */
};

/* Boundary patch kinds.  Each is named by the rho boundary type that selects
 * it in dugksFoam (the DF-side type follows bcMap, discreteVelocity.C:274-281). */
enum dugks_patch_type {
    DUGKS_PATCH_WALL = 1,      /* rho calculatedMaxwell, U/T fixedValue: diffuse Maxwell wall */
    DUGKS_PATCH_CYCLIC = 2,    /* OpenFOAM cyclic (coupled) patch pair                      */
    DUGKS_PATCH_SYMMETRY = 3,  /* rho symmetryMod (gSurf DVMsymmetry): axis-aligned mirror
                                  plane; incoming values = the mirrored velocity's outgoing
                                  value (fvDVM.C:375-454, discreteVelocity.C:733-817)       */
    DUGKS_PATCH_FARFIELD = 4,  /* rho farField, T fixedValue (gSurf farField): incoming
                                  values rho_b M(U of the owner cell, T_b)
                                  (discreteVelocity.C:574-604)                               */
    DUGKS_PATCH_FIXED = 5,     /* rho fixedValue (gSurf mixed): incoming values frozen at the
                                  boundary Maxwellian rho_b M(U_b, T_b)
                                  (discreteVelocity.C:312-344, 556-573, 1046-1060)           */
    DUGKS_PATCH_PRESSURE_IN = 6,  /* rho pressureIn: a far field whose rho_b = p / (R T_b) is
                                  set by dugks_init_equilibrium and after every step
                                  (fvDVM.C:743-770, the constructor's call 1071-1072)         */
    DUGKS_PATCH_PRESSURE_OUT = 7, /* rho pressureOut: a far field whose rho_b and T_b follow the
                                  boundary cells' characteristics, set from the initial state
                                  by dugks_init_equilibrium and after every step
                                  (fvDVM.C:772-804, 1071-1072)                                */
    DUGKS_PATCH_ZEROGRADIENT = 8, /* rho zeroGradient (gSurf zeroGradient): every face value is
                                  the owner cell's gBarPvol (discreteVelocity.C:551-555); only
                                  the outgoing half is relaxed in S7 (its field operators are
                                  no-ops, zeroGradientFvsPatchField.H:163-181)                */
    DUGKS_PATCH_SYMMETRY_PLANE = 9 /* OpenFOAM symmetryPlane patch (rho, U, T symmetryPlane; the
                                  serial symmetry boundary of doc/usage.tex:77-81): axis-aligned
                                  mirror plane.  gBarPvol takes the constraint value phi_b =
                                  phi_P (no lagged gradient row; the face still enters the
                                  least-squares geometry); outgoing values (xi.S > -VSMALL) are
                                  reconstructed, incoming ones are the mirrored velocity's face
                                  value (discreteVelocity.C:673-689, 791-815); S7 relaxes every
                                  value through the field operation and the outgoing ones
                                  (xi.S > 0) a second time (discreteVelocity.C:880-881, 905-920) */
};

/* Storage precision of the resident distributions. */
enum dugks_storage {
    DUGKS_STORAGE_FP64 = 0,    /* plain FP64 path                                 */
    DUGKS_STORAGE_FP32 = 1     /* g/h stored in FP32; equilibrium, moments in FP64 */
};

typedef struct dugks_patch {
    int32_t type;              /* dugks_patch_type                                   */
    int32_t start;             /* first face, global face numbering                  */
    int32_t size;              /* number of faces                                    */
    int32_t neighbour_patch;   /* cyclic partner patch index, else -1                */
    const double* U;           /* [size][3] boundary U (wall, fixed), else may be NULL          */
    const double* T;           /* [size]    boundary T (wall, far field, fixed), else NULL      */
    const double* rho;         /* [size]    boundary rho (far field, fixed), else NULL          */
    double pressure;           /* pressureIn / pressureOut: the patch pressure (pIn.C:72)       */
} dugks_patch;

typedef struct dugks_case {
    int32_t abi_version;       /* = DUGKS_ABI_VERSION                                */
    int32_t dim;               /* mesh.nSolutionD(): 1, 2 or 3                       */
    int32_t n_cells;
    int32_t n_internal_faces;
    int32_t n_faces;           /* internal + faces of non-empty patches              */
    const int32_t* owner;      /* [n_faces]  fvMesh::owner() / faceCells             */
    const int32_t* neighbour;  /* [n_internal_faces]                                 */
    const double* C;           /* [n_cells][3] cell centres                          */
    const double* V;           /* [n_cells]    cell volumes                          */
    const double* Sf;          /* [n_faces][3] face area vectors                     */
    const double* Cf;          /* [n_faces][3] face centres                          */
    const double* weights;     /* [n_faces] surfaceInterpolation::weights()          */
    const double* delta;       /* [n_faces - n_internal_faces][3] fvPatch::delta()   */
    int32_t n_patches;
    const dugks_patch* patches;
    /* velocity space (DVMProperties fvDVMparas + constant/Xis, constant/weights) */
    int32_t n_xi;              /* nDV, points per axis                               */
    const double* xi;          /* [n_xi] */
    const double* xi_weights;  /* [n_xi] */
    double xi_max;             /* xiMax, used by the Courant number                  */
    /* gas (DVMProperties gasProperties) */
    double R, omega, T_ref, mu_ref, Pr;
    int32_t K;                 /* KInner                                             */
    int32_t storage;           /* dugks_storage                                      */
    int32_t device;            /* HIP device ordinal                                 */
    /* Independent cases advanced together (batching).  n_cases <= 1 means one
     * case.  cell_case[c] names the case of cell c; the mesh is the disjoint
     * union of the cases' meshes (no face joins two cases). */
    int32_t n_cases;
    const int32_t* cell_case;  /* [n_cells] or NULL                                  */
    const double* mu_ref_case; /* [n_cases] per-case muRef, or NULL (= mu_ref)       */
} dugks_case;

typedef struct dugks_engine dugks_engine;

/* Builds the engine: validates the case, derives the least-squares vectors,
 * uploads mesh and velocity set once, allocates resident state.
 * Replaces the fvDVM constructor (fvDVM.C:886-1075). */
int dugks_create(const dugks_case* cs, dugks_engine** out);

/* Sets rho, U, T in every cell (q = 0) and the distributions to the Shakhov
 * equilibrium with q = 0 (discreteVelocity::initDFtoEq, discreteVelocity.C:220-249;
 * wall inComingByRho, fvDVM.C:263-309; tau, fvDVM.C:1073; Usurf, fvDVM.C:1074). */
int dugks_init_equilibrium(dugks_engine* h, const double* rho, const double* U, const double* T);

/* Advances nsteps time steps of size dt: fvDVM::evolution() (fvDVM.C:1086-1108)
 * called nsteps times with a fixed deltaT. */
int dugks_step(dugks_engine* h, double dt, int32_t nsteps);

/* As dugks_step with one time step per case: dt_case[n_cases]. */
int dugks_step_cases(dugks_engine* h, const double* dt_case, int32_t nsteps);

/* Cell fields after the last step: rho[n_cells], U[n_cells][3], T[n_cells],
 * q[n_cells][3] (fvDVM::rhoVol/Uvol/Tvol/qVol, fvDVM.C:597-728).  NULL skips. */
int dugks_get_macro(dugks_engine* h, double* rho, double* U, double* T, double* q);

/* Face fields of the last step (fvDVM.C:456-582): rho, U[3], T, q[3] on all
 * n_faces faces.  NULL skips. */
int dugks_get_face_macro(dugks_engine* h, double* rho, double* U, double* T, double* q);

/* Wall data of patch `patch` after the last step: rho_w (calculatedMaxwell value,
 * calculatedMaxwellFvPatchField.C:139-164), qWall[3], stressWall[9] (fvDVM.C:539-581).
 * Arrays are sized by the patch; NULL skips. */
int dugks_get_wall(dugks_engine* h, int32_t patch, double* rho_w, double* q_wall, double* stress_wall);

/* Courant number of fvDVM::getCoNum (fvDVM.C:1111-1119) for time step dt. */
int dugks_get_courant(dugks_engine* h, double dt, double* co_max, double* co_mean);

/* Distribution functions g~ and h~ of n cells (cells[n]) for every velocity in
 * reference order: g[n][nv], h[n][nv] (fvDVM::writeDFonCell, fvDVM.C:820-875).
 * Either output may be NULL. */
int dugks_get_df(dugks_engine* h, int32_t n, const int32_t* cells, double* g, double* h_out);

/* As dugks_get_df, but allocates nothing: no hipMalloc, hipFree or other device or host
 * allocation on its success path.  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight time loops that dump distributions every
 * step, harnesses that forbid allocation).  The cell list and the g, h rows pass through a fixed
 * device array of DUGKS_DF_NOALLOC_MAX_DOUBLES doubles (8 MiB) per GPU, held until the process
 * ends: a call needs 2 n nv + 32 ceil(n / 64) <= DUGKS_DF_NOALLOC_MAX_DOUBLES (e.g. n <= 37 cells
 * at nv = 24^3, n <= 311 at nv = 41^2), else DUGKS_ERR_INVALID.  Every call on one GPU uses
 * the same array: calls that overlap on one GPU (other threads, other engines on that GPU)
 * race and must be serialised by the caller. */
#define DUGKS_DF_NOALLOC_MAX_DOUBLES (1u << 20)
int dugks_get_df_noalloc(dugks_engine* h, int32_t n, const int32_t* cells, double* g, double* h_out);

/* Number of velocities, cells and faces. */
int dugks_get_sizes(dugks_engine* h, int32_t* nv, int32_t* n_cells, int32_t* n_faces);

/* Step-loop graphs (no reference counterpart): with graphs on (the default) dugks_step
 * with nsteps >= 2 replays a captured HIP graph of two steps instead of launching each
 * kernel; results are identical either way.  Profiling runs eagerly. */
int dugks_set_graphs(dugks_engine* h, int32_t on);

/* Online launch-shape tuning (no reference counterpart): with autotune on, the next steps
 * run one candidate work decomposition of the face kernel and of the cell kernel each (after
 * one warm-up step; every candidate in two rounds), timed with HIP events, then keep the one
 * with the shortest time of each.  The steps are
 * real steps of the simulation; the candidates differ only in how velocity sums are split,
 * i.e. in round-off.  dugks_get_launch / dugks_get_cell_launch report the shapes in use. */
int dugks_set_autotune(dugks_engine* h, int32_t on);
/* Steps still to run before the tuning started by dugks_set_autotune ends (0 when idle). */
int dugks_autotune_steps_left(dugks_engine* h, int32_t* steps);
int dugks_get_cell_launch(dugks_engine* h, int32_t* nt, int32_t* nib, int32_t* n_blocks);

/* Face-value mode (no reference counterpart): with it on, the face kernel stores every face
 * value it rebuilds ([face][velocity], storage precision) and the cell kernel reads them instead
 * of rebuilding them from gbar and its gradients.  FP64 results are bit-identical either way;
 * with FP32 storage the face values are rounded to FP32 like the other stored rows.  Costs
 * n_faces x nv_padded x (4 or 8) bytes per carried distribution.  On by default when those
 * buffers take at most half of the free device memory at create. */
int dugks_set_face_values(dugks_engine* h, int32_t on);
int dugks_get_face_values(dugks_engine* h, int32_t* on);

/* Work decomposition of the face kernel (no reference counterpart; for measurement and
 * tuning): workgroup size, velocity iterations per thread in a block, number of
 * velocity blocks and the padded velocity count nt * nib * n_blocks.  NULL skips. */
int dugks_get_launch(dugks_engine* h, int32_t* nt, int32_t* nib, int32_t* n_blocks, int32_t* nv_padded);

/* Complete resident state as an opaque byte image: resuming from it continues
 * bit-identically.  dugks_state_size gives the byte count. */
int dugks_state_size(dugks_engine* h, int64_t* bytes);
int dugks_save_state(dugks_engine* h, void* buf, int64_t bytes);
int dugks_load_state(dugks_engine* h, const void* buf, int64_t bytes);

/* Synchronises the stream and reports the device time of the kernels of the
 * last dugks_step call (ms), measured with HIP events on the engine stream. */
int dugks_last_step_ms(dugks_engine* h, double* ms);

/* Per-kernel timing (no reference counterpart; for measurement).  With
 * profiling on, every step records HIP events on the engine stream around its
 * kernels; dugks_get_profile returns in ms[6] the accumulated device time of
 * face pass 1 (k_face: S2 reconstruction + S6 face-moment sums; in face-value
 * mode k_face_values: S2 face values, cell-wise), face pass 2 (k_face_moments:
 * S6 face-moment sums of the stored values; 0 outside face-value mode),
 * k_face_fin (S3 wall density, face macros), k_cell (S7 relaxation, S8 update,
 * S9 moment sums), k_cell_fin (S9 cell macros) and k_cell_s1 (S1 of the next
 * step), and the number of steps since profiling was switched on. */
int dugks_set_profiling(dugks_engine* h, int32_t on);
int dugks_get_profile(dugks_engine* h, double* ms, int64_t* steps);

void dugks_destroy(dugks_engine* h);
const char* dugks_last_error(dugks_engine* h);
int dugks_abi_version(void);

/* ---- Discrete adjoint (no reference counterpart) ----------------------------------------
 * Gradient of a time-averaged mass-flux objective with respect to the wall temperatures:
 *
 *   J = sum over steps n = first_step .. n_steps of  sum_f c_f (rho U)_f^n . S_f,
 *
 * (rho U)_f^n = sum_k w_k xi_k g_f,k the face momentum of step n (the S6 face moments of
 * fvDVM.C:473-483, i.e. dugks_get_face_macro's rho * U after that step), c_f = face_coef[f].
 * dJ_dTw[b] = dJ/dT of boundary face n_internal_faces + b: the wall temperature of a wall face,
 * the boundary temperature of a fixed, far-field or pressureIn face, the patch temperature of a
 * pressureOut face (it enters the first step only), 0 on cyclic, zeroGradient and symmetry faces.
 * The run starts from the Shakhov equilibrium of rho, U, T with q = 0 (as
 * dugks_init_equilibrium) and advances n_steps FP64 steps of size dt, keeping a checkpoint every
 * checkpoint_every steps (0: ceil(sqrt(n_steps))); the reverse sweep re-runs each segment.
 * Cases: every patch kind; batches (n_cases > 1) with their per-case muRef and the one dt; FP64 (the case's
 * storage field is ignored).  Device memory: about
 * (ceil(n_steps / C) + C + 1) x (2 n_cells + 2 n_boundary_faces) x nv x 8 bytes for the trajectory
 * plus about (8 + 4 dim) n_cells x nv x 8 and 4 n_faces x nv x 8 bytes of work arrays;
 * DUGKS_ERR_MEMORY when that does not fit.  Where it also fits, a face tape of
 * (C - 1) x (2 nv + 32) x n_faces x 8 bytes spares the reverse steps a recomputation (results are
 * identical without it). */
typedef struct dugks_adjoint dugks_adjoint;
/* *out is set also when the call fails, so that dugks_adjoint_last_error can explain it;
 * release it with dugks_adjoint_destroy either way. */
int dugks_adjoint_create(const dugks_case* cs, dugks_adjoint** out);
int dugks_adjoint_run(dugks_adjoint* a, const double* rho, const double* U, const double* T, double dt,
                      int32_t n_steps, int32_t first_step, int32_t checkpoint_every,
                      const double* face_coef, double* J, double* dJ_dTw);
/* Cell fields rho, U[3], T, q[3] after the last run's final step.  NULL skips. */
int dugks_adjoint_get_macro(dugks_adjoint* a, double* rho, double* U, double* T, double* q);
/* Device time of the last run's forward sweep and of its reverse sweep (with re-runs), ms. */
int dugks_adjoint_last_ms(dugks_adjoint* a, double* forward_ms, double* reverse_ms);
void dugks_adjoint_destroy(dugks_adjoint* a);
const char* dugks_adjoint_last_error(dugks_adjoint* a);

#ifdef __cplusplus
}
#endif

#endif /* DUGKS_H */
