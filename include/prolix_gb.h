/*
 * prolix_gb.h -- C ABI of the MI355X (gfx950) GB-OBC2 implicit-solvent MD engine.
 *
 * Plain C types only: host pointers for topology/parameters, device pointers
 * for per-step arrays, and an opaque `void*` that is a hipStream_t (NULL =
 * the device's legacy default stream). Captured step graphs replay on the
 * engine's own stream, ordered after and before the caller's stream by events.
 * Every function returns PRX_OK (0) or a PRX_ERR_* code; prx_last_error()
 * returns a thread-local message for the last failure on the calling thread.
 * No C++ exception crosses this boundary.
 *
 * Each entry point names the prolix interface it replaces (paths under
 * prolix's src/prolix/).
 *
 * Units: A, kcal/mol, e, amu; time in fs / ps at the boundary (AKMA inside).
 * Threading: a handle is used from one host thread at a time.
 */
#ifndef PROLIX_GB_H
#define PROLIX_GB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRX_ABI_VERSION 1

enum prx_status {
  PRX_OK = 0,
  PRX_ERR_INVALID = 1,  /* bad argument, shape or index */
  PRX_ERR_HIP = 2,      /* a HIP runtime call failed */
  PRX_ERR_CAPACITY = 3, /* pool capacity / HBM budget exhausted */
  PRX_ERR_STATE = 4     /* call out of order (e.g. steps before init) */
};

/* Per-system energy outputs, the terms summed by total_energy
 * (physics/system.py:229-248). energies[] rows are PRX_N_ENERGY doubles. */
enum prx_energy_term {
  PRX_E_BOND = 0,
  PRX_E_ANGLE = 1,
  PRX_E_UREY_BRADLEY = 2,
  PRX_E_DIHEDRAL = 3,
  PRX_E_IMPROPER = 4,
  PRX_E_CMAP = 5,
  PRX_E_LJ = 6,
  PRX_E_EXCEPTION = 7,
  PRX_E_GB = 8,
  PRX_E_COULOMB = 9,
  PRX_E_ACE = 10,
  PRX_E_TOTAL = 11,
  PRX_N_ENERGY = 12
};

/* One system's topology and parameters (host arrays, read during the call).
 * Mirrors PhysicsSystem (typing.py:248-310) plus the make_energy_fn options
 * (physics/system.py:29-37, 110-135). NULL pointers with a zero count are
 * fine for every optional term. */
typedef struct prx_system_desc {
  int32_t n_atoms;
  const double* charges;      /* (n) e */
  const double* sigmas;       /* (n) A, clamped to >= 1e-6 (system.py:107) */
  const double* epsilons;     /* (n) kcal/mol */
  const double* radii;        /* (n) A, OBC intrinsic radii */
  const double* scaled_radii; /* (n) A absolute descreening radii; NULL -> radii */
  const double* masses;       /* (n) amu; NULL -> 1 */

  int32_t n_bonds;            /* harmonic 1/2 k (r - r0)^2, params [r0, k] */
  const int32_t* bond_idx;    /* (n_bonds, 2) */
  const double* bond_params;  /* (n_bonds, 2) */
  int32_t n_ub;               /* Urey-Bradley, same form as bonds */
  const int32_t* ub_idx;
  const double* ub_params;
  int32_t n_angles;           /* 1/2 k (theta - theta0)^2, params [theta0, k] */
  const int32_t* angle_idx;   /* (n_angles, 3) */
  const double* angle_params; /* (n_angles, 2) */
  int32_t n_dihedrals;        /* sum_t k (1 + cos(n phi - phase)), params [n, phase, k] */
  int32_t dihedral_terms;     /* T */
  const int32_t* dihedral_idx;   /* (n_dihedrals, 4) */
  const double* dihedral_params; /* (n_dihedrals, T, 3) */
  int32_t n_impropers;
  int32_t improper_terms;        /* T */
  int32_t improper_param_width;  /* 3: periodic [n, phase, k]; 2: harmonic [k, phi0] */
  const int32_t* improper_idx;   /* (n_impropers, 4) */
  const double* improper_params; /* (n_impropers, T, width) */
  int32_t n_cmap;
  const int32_t* cmap_torsions;  /* (n_cmap, 5): phi = atoms 0..3, psi = atoms 1..4 */
  const int32_t* cmap_map;       /* (n_cmap) map index */
  int32_t n_cmap_maps;
  int32_t cmap_grid;             /* G (24) */
  const double* cmap_grids;      /* (n_cmap_maps, G, G) raw energies stored [psi, phi] */

  int32_t excl_cols;             /* C of the (n, C) exclusion rows (0 = no exclusions) */
  const int32_t* excl_idx;       /* (n, C) partner index, -1 = empty */
  const double* excl_scale_vdw;  /* (n, C) */
  const double* excl_scale_elec; /* (n, C) */

  int32_t n_exceptions;          /* explicit 1-4 pairs (bonded.py:188-228) */
  const int32_t* exc_pairs;      /* (n_exceptions, 2) */
  const double* exc_sigma;
  const double* exc_epsilon;
  const double* exc_chargeprod;

  double cutoff;              /* A; 0 = no cutoff. LJ/Coulomb only unless gb_cutoff */
  double coulomb_constant;    /* 332.0637 (system.py:135) */
  double gb_coulomb_constant; /* constant of the GB prefactor */
  double solute_dielectric;   /* 1.0 */
  double solvent_dielectric;  /* 78.5 */
  double dielectric_offset;   /* 0.09 A */
  int32_t implicit_solvent;   /* 1: GB-OBC2 + ACE on */
  int32_t gb_cutoff;          /* 1 (with cutoff > 0): the Born sums and the GB pair terms
                                 also run over the pairs closer than the cutoff only --
                                 the neighbour-list GB of generalized_born.py:266-415 as
                                 the batched path calls it (batched_energy.py:999-1010),
                                 with the list taken as every pair within the cutoff:
                                 offset radii >= 1e-4, 1/B >= 1e-4, self term q^2/B.
                                 0: GB over all pairs (physics/system.py:196-209). */
} prx_system_desc;

/* Langevin BAOAB parameters for one system (jax_md nvt_langevin as used at
 * physics/simulate.py:245-255; AKMA conversion as simulate.py:770-772). */
typedef struct prx_langevin_params {
  double dt_fs;        /* time step, fs */
  double temperature;  /* K */
  double gamma_ps;     /* friction, 1/ps (0 = velocity Verlet) */
  uint64_t seed;       /* Philox4x32-10 key of this system's noise stream */
  int32_t safety_caps; /* 1: the batched step's caps (batched_simulate.py:269-358):
                          |p/m| <= 20 A/tau, |dx| <= 0.5 A per step, |F| <= 1e4
                          kcal/mol/A, non-finite forces -> 0. 0: plain nvt_langevin */
  int32_t reserved;
} prx_langevin_params;

typedef struct prx_engine prx_engine;

/* ---- library ---------------------------------------------------------- */

/* Message for the last failed call on this thread ("" if none). */
const char* prx_last_error(void);
int32_t prx_abi_version(void);

/* ---- systems: make_energy_fn's set-up (physics/system.py:29-171) --------
 * Upload n_systems systems packed without padding into one engine. The
 * engine owns every device buffer until prx_engine_destroy. */
int prx_system_create(const prx_system_desc* descs, int32_t n_systems, prx_engine** out);
int prx_engine_destroy(prx_engine* eng);

/* Packed layout: offsets[s] = first atom of system s, offsets[n] = total.
 * The atom order inside a system is the caller's. */
int prx_engine_layout(const prx_engine* eng, int32_t* atom_offsets, int32_t capacity);

/* ---- energy / forces: total_energy(r) and -jax.grad of it -------------
 * Replaces make_energy_fn's total_energy (physics/system.py:229-248) and
 * its gradient (custom_vjp bwd of optimization.py:75-126, 250-297 plus the
 * Born-radius chain rule, generalized_born.py:651-817).
 * positions/forces: device (total_atoms, 3) float32 in packed order.
 * energies: device (n_systems, PRX_N_ENERGY) float64, or NULL.
 * born_radii: device (total_atoms) float32, or NULL. Stream-ordered. */
int prx_energy_forces(prx_engine* eng, const float* positions, float* forces, double* energies,
                      float* born_radii, void* stream);

/* Allocation-free variant of prx_energy_forces for one system given by its
 * description, with no engine: every device buffer it needs (parameters,
 * topology, work list, partial sums) is carved from a static device array of
 * 32 MiB that each GPU holds from library load until the process ends; the
 * call makes no hipMalloc / hipFree and creates no stream. No counterpart in
 * prolix: it is for callers that cannot allocate device memory (code inside a
 * serving loop or harness that forbids allocation, or a stateless custom-call
 * handler that keeps no engine between calls).
 * Limits: n_atoms <= 8192, and the system's buffers must fit the 32 MiB
 * (PRX_ERR_CAPACITY otherwise; an 8,192-atom protein needs about 24 MiB).
 * positions / forces / energies / born_radii as for prx_energy_forces (one
 * system: energies is one row). The topology is read from host memory and
 * uploaded with synchronous copies; the call returns after its results are
 * written. All calls on one GPU share the array: calls that overlap in time
 * on one GPU (other host threads, other streams) race and must not be made. */
int prx_energy_forces_noalloc(const prx_system_desc* desc, const float* positions, float* forces,
                              double* energies, float* born_radii, void* stream);

/* ---- integrator: nvt_langevin init_fn / apply_fn ------------------------
 * prx_integrator_init = init_fn(key, R, mass) (settle.py:1676-1697):
 * positions (device, packed) become the state; momenta (device, packed) or
 * NULL to draw Maxwell-Boltzmann momenta from each system's Philox stream.
 * params: host array with one entry per system. Computes the initial forces. */
int prx_integrator_init(prx_engine* eng, const float* positions, const float* momenta,
                        const prx_langevin_params* params, void* stream);
/* n_steps BAOAB steps of every system (settle.py:1699-1744, jax_md
 * nvt_langevin apply_fn), replayed from a captured hipGraph. */
int prx_integrator_steps(prx_engine* eng, int32_t n_steps, void* stream);
/* Replace the atoms' masses (host, packed, amu) before prx_integrator_init;
 * nvt_langevin's `mass` argument (settle.py:1676-1697). */
int prx_engine_set_masses(prx_engine* eng, const double* masses_amu);
/* Copy state out (device pointers, packed; any may be NULL). */
int prx_integrator_state(prx_engine* eng, float* positions, float* momenta, float* forces,
                         void* stream);
/* Per-system kinetic energy sum p^2 / 2m (kcal/mol), float64 reduction. */
int prx_integrator_kinetic(prx_engine* eng, double* kinetic, void* stream);
/* Steps taken so far by system s (host value, synchronises the stream). */
int prx_integrator_step_count(prx_engine* eng, int32_t system, int64_t* steps);
/* Subtract each system's mass-weighted centre of mass from its positions
 * (float64 reduction); momenta and stored forces are unchanged. Replaces
 * recenter_com (batched_simulate.py:58-81) as applied by
 * apply_com_correction (batched_simulate.py:122-155) during streaming
 * production (batched_simulate.py:1226-1422). */
int prx_recenter_com(prx_engine* eng, void* stream);
/* Safety-cap event counters of system s (atom-steps that hit the velocity,
 * force, constraint and displacement caps; warn_counts of typing.py:558-563).
 * Host out, 4 values; synchronises the stream. */
int prx_integrator_warnings(prx_engine* eng, int32_t system, int64_t* counts);
/* Noise-stream position of each system (host array, one entry per system in
 * packed order, each in [0, 2^32)): the number of BAOAB steps it has taken.
 * Called after prx_integrator_init with the momenta of a saved state, it
 * resumes that state's Langevin trajectory bit for bit (the step counter is
 * the Philox counter), as the resumed production of
 * batched_simulate.py:1300-1305 continues from a saved state and step. */
int prx_integrator_set_step_counts(prx_engine* eng, const int64_t* steps, void* stream);
/* Explosion check (simulate.py:1050-1071, batched_simulate.py:158-205) kept
 * on the device: nonfinite receives, per system in packed order (host array),
 * the number of atom-steps since its prx_integrator_init / admission that
 * ended with a non-finite position or momentum. Synchronises the stream. */
int prx_integrator_health(prx_engine* eng, int64_t* nonfinite, void* stream);
/* Verlet skin (A, default 0) of the neighbour lists of the engine's gb_cut
 * systems (cutoff > 0 with gb_cutoff = 1), which take their pair work from a
 * list built on the device: per 64-atom tile, the atoms of later tiles
 * within cutoff + skin of its bounding box. With skin = 0 the list is rebuilt
 * at every evaluation; with skin > 0 when an atom has moved more than skin / 2
 * since the last build (and at every eager evaluation and graph launch).
 * Results do not depend on the skin: every pair is tested against the cutoff.
 * Replaces make_neighbor_list_fn's jax_md list (physics/neighbor_list.py:155-172,
 * dr_threshold) and its rebuild rule. */
int prx_engine_set_nl_skin(prx_engine* eng, double skin);
/* How prx_integrator_steps runs the steps. Every path runs the same
 * per-item and per-tile device code on the same work lists, so the
 * trajectory is the same bit for bit:
 * 1 = captured HIP graphs of the per-pass kernels;
 * 2 = one persistent cooperative kernel per 1,000 steps whose passes are
 *     separated by grid-wide barriers (the call returns after the steps);
 * 3 = fused: one kernel per 1,000 steps in which one workgroup runs all the
 *     passes of one system with workgroup barriers only, systems side by
 *     side; PRX_ERR_STATE if a system uses a neighbour list;
 * 0 = auto (default): fused when every system has at most 192 atoms and
 *     none uses a neighbour list (measured fastest there on MI355X), graphs
 *     otherwise (fastest from 256 to 2,500 atoms).
 * PRX_ERR_STATE for 2 where cooperative launches are unavailable. No
 * counterpart in prolix (its loop is jax.lax.fori_loop,
 * physics/simulate.py:253-254). */
int prx_engine_set_step_mode(prx_engine* eng, int32_t mode);
/* Run the bonded kernel on a second stream beside the pair passes (a
 * parallel branch of the step graph) when the engine holds at least
 * min_atoms atoms; below it, in line. Default 100,000. Results are the same
 * either way, bit for bit. No counterpart in prolix. */
int prx_engine_set_fork_min_atoms(prx_engine* eng, int64_t min_atoms);

/* ---- measurement --------------------------------------------------------
 * Launch each force pass of one step n_reps times back to back, the run
 * bracketed by one pair of HIP events on the engine stream, at the current
 * positions; ms receives the mean duration per launch of: Born pairs, Born
 * gather, force pairs, force gather, chain pairs, bonded, chain gather
 * (7 doubles). The integrator's
 * positions and momenta are untouched; its stored forces are recomputed at
 * the same positions.
 * No counterpart in prolix (its timing is wall-clock around jitted calls). */
int prx_profile_passes(prx_engine* eng, int32_t n_reps, double* ms);
/* Take n_steps (1..1000) integrator steps through a step graph that holds a
 * one-lane kernel storing the GPU wall clock before and after every pass; ms
 * receives each pass's mean duration inside the graph, launch gap included
 * (same 7 passes and order as prx_profile_passes).
 * The integrator advances n_steps like prx_integrator_steps.
 * No counterpart in prolix. */
int prx_profile_steps(prx_engine* eng, int32_t n_steps, double* ms);
/* Number of pair work items (64x64 tiles) of the engine's current work list. */
int prx_engine_work_items(const prx_engine* eng, int64_t* n_items);

/* ---- serving pool -------------------------------------------------------
 * A pool is an engine whose systems come and go. Capacities bound the
 * device arenas (atoms, bonded entries, CMAP coefficients, exclusion tiles);
 * prx_pool_admit fails with PRX_ERR_CAPACITY when a system does not fit.
 * No counterpart in prolix, whose batched path pads to buckets
 * (padding.py:573-654, batched_simulate.py:1125-1422). */
typedef struct prx_pool_limits {
  int64_t max_atoms;
  int64_t max_bonded;     /* bonds + UB + angles + torsion terms + CMAP + exceptions */
  int64_t max_cmap_coeff; /* floats of CMAP patch coefficients (16 per patch) */
  int64_t max_excl_tiles; /* 64x64 tiles that hold exclusions */
  int32_t max_systems;
} prx_pool_limits;

int prx_pool_create(const prx_pool_limits* limits, prx_engine** out);
/* Admit one system with its start positions (host, (n,3) float32) and
 * integrator parameters. The new system joins the next prx_integrator_steps
 * call: its momenta are drawn from its Philox stream and the forces computed
 * then (or at the next read of the pool), once for every system admitted
 * since the last step call. slot receives its id. */
int prx_pool_admit(prx_engine* pool, const prx_system_desc* desc, const float* positions_host,
                   const prx_langevin_params* params, int32_t* slot, void* stream);
/* Stop one system after limit BAOAB steps in all (0: no limit, the default
 * at admission). A system at its limit keeps its positions, momenta and step
 * count while the others step on, so a step call may run past the end of one
 * request without changing its result. Synchronises the streams. Serving
 * only; prolix runs every system of a bucket for the same step count
 * (batched_simulate.py:1125-1422). */
int prx_pool_set_step_limit(prx_engine* pool, int32_t slot, int64_t limit, void* stream);
/* Remove a system; its arena ranges are reused by later admissions. */
int prx_pool_release(prx_engine* pool, int32_t slot, void* stream);
/* Copy one system's positions / momenta / forces to host ((n,3) float32 each,
 * any may be NULL) and its step count; synchronises the stream. */
int prx_pool_read(prx_engine* pool, int32_t slot, float* positions, float* momenta, float* forces,
                  int64_t* steps, void* stream);
/* Number of live systems and their slot ids (capacity entries max). */
int prx_pool_slots(const prx_engine* pool, int32_t* slots, int32_t capacity, int32_t* n_live);
/* Energies of one live slot at its current positions (host out, PRX_N_ENERGY doubles). */
int prx_pool_energy(prx_engine* pool, int32_t slot, double* energies_host, void* stream);

/* ---- neighbour-list object -------------------------------------------------
 * The list of make_neighbor_list_fn (physics/neighbor_list.py:155-172, a
 * jax_md Dense list with mask_self): for each atom i of the (n, 3) float32
 * device positions, the atoms j != i closer than radius (cutoff + the
 * list's skin) in increasing j, written to the device (n, capacity) int32
 * idx with n after the last entry; counts (device, n int32) receives the
 * true numbers, so counts[i] > capacity flags an overflow. workspace: device
 * float array of 8 * ceil(n / 64). Stream-ordered. The engine's own pair
 * work does not use this list (gb_cut systems build their tile lists). */
int prx_neighbor_list_dense(const float* positions, int32_t n_atoms, double radius, int32_t capacity,
                            int32_t* idx, int32_t* counts, float* workspace, void* stream);

/* ---- host-side planning (no device work; usable without a GPU) --------- */

/* Bicubic CMAP patch coefficients of prolix's cmap.py:242-286 for one grid
 * stored [psi, phi]: out is (G, G, 16) float64 indexed [phi, psi]. */
int prx_cmap_coefficients(const double* grid, int32_t G, double* out);

/* Number of work items and of exclusion tiles the pair kernels would use for
 * one system in an engine of its own: the 64x64 tile pairs I <= J, or for a
 * gb_cut system its diagonal and exclusion tile pairs plus the neighbour
 * list's chunk slots. */
int prx_plan_tiles(const prx_system_desc* desc, int64_t* n_items, int64_t* n_excl_tiles);

#ifdef __cplusplus
}
#endif

#endif /* PROLIX_GB_H */
