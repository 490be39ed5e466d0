/* oxdna_hip.h — C ABI of the MI355X (gfx950) oxDNA1 / oxDNA2 / oxRNA2 force engine.
 *
 * Plain C: extern "C" functions, POD structs, pointers and sizes. No C++ or torch
 * types cross this boundary. Every function returns an int status (0 = success,
 * negative = error); the message of the last error on a context is returned by
 * oxdna_last_error(ctx), and context-free entry points report through
 * oxdna_last_error(NULL). No exception ever crosses the ABI.
 *
 * Two levels are exported:
 *   (1) stateless kernel drivers with the data layouts of the host MD engine the
 *       reference plugs into (HOOMD-blue 6.x): Scalar4 position (type id in the
 *       bits of .w), Scalar4 orientation quaternion stored (w, x, y, z), a full
 *       neighbour list (n_neigh / nlist / head_list), ForceReal4 force (energy in
 *       .w) and torque, a pitched 6-component virial. They replace the reference's
 *       GPU kernel drivers one for one (citations on each declaration).
 *   (2) a context API that owns a whole system on one GPU: state, GPU cell-list
 *       neighbour list, the fused force step, a Langevin integrator and batched
 *       replica ensembles.
 *
 * Units are oxDNA simulation units throughout.
 */
#ifndef OXDNA_HIP_H
#define OXDNA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OXDNA_ABI_VERSION 1

/* ------------------------------------------------------------------------- */
/* Term identifiers (bit positions in a term mask).                          */
/* ------------------------------------------------------------------------- */
enum oxdna_term {
    OXDNA_TERM_EXCVOL = 0,     /* nonbonded excluded volume   (EvaluatorPairOxDNAExcVol)      */
    OXDNA_TERM_HBOND = 1,      /* hydrogen bonding            (EvaluatorPairOxDNAHBond)       */
    OXDNA_TERM_CROSS = 2,      /* cross-stacking              (EvaluatorPairOxDNACrossStacking) */
    OXDNA_TERM_COAX = 3,       /* coaxial stacking            (EvaluatorPairOxDNACoaxStacking) */
    OXDNA_TERM_DEBYE = 4,      /* Debye-Hueckel               (EvaluatorPairOxDNADebye)       */
    OXDNA_TERM_HBONDCROSS = 5, /* fused HB + cross-stacking   (EvaluatorPairOxDNAHBondCross)  */
    OXDNA_TERM_BONDED = 6,     /* FENE + bonded exc. vol + stacking (OxDNABondedForceCompute) */
    OXDNA_N_TERMS = 7
};

/* Per-term energy slots, in the column order of oxDNA's split_energy.dat
 * (reference pytest/_helpers.py:27-30). The fused HBondCross term reports its two
 * parts in the hb and cr_stack slots. */
enum oxdna_energy_slot {
    OXDNA_E_FENE = 0, OXDNA_E_BEXC = 1, OXDNA_E_STACK = 2, OXDNA_E_NEXC = 3,
    OXDNA_E_HB = 4, OXDNA_E_CRSTACK = 5, OXDNA_E_CXSTACK = 6, OXDNA_E_DEBYE = 7,
    OXDNA_N_ENERGY = 8
};

/* ------------------------------------------------------------------------- */
/* Parameter structs. FP64 on the host like the reference's Scalar fields;   */
/* booleans are int32. Field order and meaning follow the reference structs. */
/* ------------------------------------------------------------------------- */

/* OxDNAExcVolParams (reference src/EvaluatorPairOxDNAExcVol.h:30-39).
 * Site pair k: 0 = back-back, 1 = base-base, 2 = base(i)-back(j), 3 = back(i)-base(j). */
typedef struct oxdna_excvol_params {
    double eps;
    double base_site[3];
    double back_site[3];
    double sigma[4];
    double rstar[4];
    double b[4];
    double rc[4];
} oxdna_excvol_params;

/* OxDNAHBondParams (reference src/EvaluatorPairOxDNAHBond.h:30-37).
 * f4 sets are (a, b, t0, ts, tc). */
typedef struct oxdna_hbond_params {
    double eps;
    double shift_factor;
    double base_site[3];
    double f1_a, f1_r0, f1_rlow, f1_rhigh, f1_rclow, f1_rchigh, f1_blow, f1_bhigh;
    double t1[5], t23[5], t4[5], t78[5];
} oxdna_hbond_params;

/* OxDNACrossStackingParams (reference src/EvaluatorPairOxDNACrossStacking.h:30-36). */
typedef struct oxdna_crossstacking_params {
    double base_site[3];
    double f2_k, f2_r0, f2_rc, f2_rlow, f2_rhigh, f2_rclow, f2_rchigh, f2_blow, f2_bhigh;
    double t1[5], t23[5], t4[5], t78[5];
    int32_t t4_enabled; /* 1: keep f4(theta4) (oxDNA); 0: drop it (oxRNA) */
    int32_t _pad;
} oxdna_crossstacking_params;

/* OxDNAHBondCrossParams (reference src/EvaluatorPairOxDNAHBondCross.h:37-47), plus
 * cs_t4_enabled, which the reference lacks, so the fused form can express the oxRNA
 * cross-stacking (theta4 dropped). Default 1 reproduces the reference exactly. */
typedef struct oxdna_hbondcross_params {
    double base_site[3];
    double eps;
    double shift_factor;
    double f1_a, f1_r0, f1_rlow, f1_rhigh, f1_rclow, f1_rchigh, f1_blow, f1_bhigh;
    double hb_t1[5], hb_t23[5], hb_t4[5], hb_t78[5];
    double f2_k, f2_r0, f2_rc, f2_rlow, f2_rhigh, f2_rclow, f2_rchigh, f2_blow, f2_bhigh;
    double cs_t1[5], cs_t23[5], cs_t4[5], cs_t78[5];
    int32_t cs_t4_enabled;
    int32_t _pad;
} oxdna_hbondcross_params;

/* OxDNACoaxStackingParams (reference src/EvaluatorPairOxDNACoaxStacking.h:32-47).
 * t1_mode 0 = oxDNA1 reflection f4(t)+f4(2pi-t); 1 = oxDNA2 f4(t)+SA(t-SB)^2. */
typedef struct oxdna_coaxstacking_params {
    double stack_site[3];
    double stack_back_ref[3];
    double gamma;
    double f2_k, f2_r0, f2_rc, f2_rlow, f2_rhigh, f2_rclow, f2_rchigh, f2_blow, f2_bhigh;
    double t1[5], t4[5], t56[5];
    double phi3[4]; /* f5 (a, b, xc, xs) */
    double phi4[4];
    int32_t t1_mode;
    int32_t phi3_enabled;
    double t1_sa, t1_sb;
    int32_t rna_coax;
    int32_t _pad;
} oxdna_coaxstacking_params;

/* OxDNADebyeParams (reference src/EvaluatorPairOxDNADebye.h:28-31). */
typedef struct oxdna_debye_params {
    double back_site[3];
    double minus_kappa, prefactor, b_smooth, rc, rhigh;
} oxdna_debye_params;

/* oxdna_bonded_params (reference src/OxDNABondedParams.h:35-73). Bonded exc. vol
 * site pair k: 0 = base-base, 1 = base(A)-back(B), 2 = back(A)-base(B). st_eps is
 * indexed [type of A (3' end)][type of B (5' neighbour)], row-major. */
typedef struct oxdna_bonded_params {
    double eps_backbone, r0_backbone, delta_backbone, delta2;
    double back_site[3];
    double base_site[3];
    double excl_eps;
    double bx_sigma[3], bx_rstar[3], bx_b[3], bx_rc[3];
    int32_t stacking_enabled;
    int32_t rna_stacking;
    double stack_site[3];
    double stack_back_ref[3];
    double st_gamma;
    double st_a, st_r0, st_rlow, st_rhigh, st_rclow, st_rchigh, st_blow, st_bhigh, st_shift_factor;
    double st_eps[16];
    double t4[5]; /* f4 theta4 (a, b, t0, ts, tc) */
    double t5[5]; /* f4 theta5 == theta6 */
    double p1[4]; /* f5 phi1 == phi2 (a, b, xc, xs) */
    double stack_3_site[3], stack_5_site[3];
    double bbvector_3[3], bbvector_5[3];
    double tB1[5], tB2[5];
} oxdna_bonded_params;

/* Orthorhombic periodic box (the reference's systems are cubic). */
typedef struct oxdna_box {
    double L[3];
} oxdna_box;

/* ------------------------------------------------------------------------- */
/* Errors and library information.                                            */
/* ------------------------------------------------------------------------- */
typedef struct oxdna_ctx oxdna_ctx;

/* Message of the last failed call on ctx (or of a context-free call when ctx is
 * NULL). Never NULL; "" when there was no error. */
const char* oxdna_last_error(const oxdna_ctx* ctx);
int oxdna_abi_version(void);
/* Name of the device the library runs on (e.g. "gfx950"), written into buf. */
int oxdna_device_name(int device, char* buf, size_t buflen);

/* ------------------------------------------------------------------------- */
/* (1) Stateless kernel drivers, HOOMD layouts, device pointers, one stream.  */
/* Each call enqueues its launches on `stream` and takes its device workspace  */
/* (parameter block, frames, energy partials) as a stream-ordered allocation   */
/* from a memory pool the library keeps per device, returned to the pool on    */
/* the same stream: calls share no device memory, so calls on one device from */
/* several streams or threads may overlap. Errors of these context-free calls  */
/* are reported per thread (oxdna_last_error(NULL)).                           */
/* ------------------------------------------------------------------------- */

/* Nonbonded pair forces for one term over a full neighbour list. Replaces the
 * reference's per-term instantiation of HOOMD's gpu_compute_pair_aniso_forces<E>
 * (src/AnisoPotentialPairOxDNA{ExcVol,HBond,CrossStacking,HBondCross,CoaxStacking,
 * Debye}GPUKernel.cu:14-18) driven by AnisoPotentialPairGPU<E>::computeForces.
 *   params:  ntypes*ntypes table of the term's struct above (row-major [ti][tj]),
 *            any field may differ between type pairs (each pair uses its own struct);
 *            ntypes is 1..4 (the oxDNA type set A, G, C, T/U)
 *   rcutsq:  ntypes*ntypes table of squared COM cut-offs (host memory)
 *   d_charge may be NULL for every term except OXDNA_TERM_DEBYE.
 *   Outputs are overwritten: force.xyz on i, force.w = half the pair energies of
 *   i, torque.xyz, and (if d_virial != NULL) virial[k*virial_pitch + i] =
 *   1/2 sum_j dr_ij (x) F_ij for k in xx,xy,xz,yy,yz,zz.
 *   d_term_energy (optional, device, 8 doubles) receives the term's total energy
 *   in its split_energy slots (overwritten). stream may be NULL (default stream). */
int oxdna_pair_forces(int term,
                      float* d_force, float* d_torque, float* d_virial, size_t virial_pitch,
                      unsigned int N,
                      const double* d_pos, const double* d_orientation, const double* d_charge,
                      const oxdna_box* box,
                      const unsigned int* d_n_neigh, const unsigned int* d_nlist,
                      const size_t* d_head_list,
                      int ntypes, const void* params, const double* rcutsq,
                      double* d_term_energy, void* stream);

/* oxdna_pair_forces with no device allocation: the same arguments, kernels and
 * outputs (bit-identical), but its scratch (parameter block, frames, energy
 * partials) is one fixed device array inside the library instead of memory it
 * allocates, and the parameter block reaches the device as kernel arguments (no
 * host-to-device copy). No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Limit: N <= 65536 particles (larger N fails with
 * an error); the array (3.3 MiB) stays on each device that ran the call until the
 * process ends. Every call on one device shares that array, so calls that overlap
 * on one device (other streams or threads) race. */
int oxdna_pair_forces_noalloc(int term,
                              float* d_force, float* d_torque, float* d_virial, size_t virial_pitch,
                              unsigned int N,
                              const double* d_pos, const double* d_orientation, const double* d_charge,
                              const oxdna_box* box,
                              const unsigned int* d_n_neigh, const unsigned int* d_nlist,
                              const size_t* d_head_list,
                              int ntypes, const void* params, const double* rcutsq,
                              double* d_term_energy, void* stream);

/* Bonded force (FENE in FP64 + bonded exc. vol + oxDNA/oxRNA stacking). Replaces
 * gpu_compute_oxdna_bonded_forces (reference src/OxDNABondedGPU.cuh:25-40,
 * src/OxDNABondedGPU.cu:125-159). Bond table in HOOMD's GPU layout: for bond slot
 * b of particle i, d_blist[2*(b*pitch+i)] = partner index, [..+1] = bond type;
 * d_bpos[b*pitch+i] = 0 if i is A (3' end of the bond), 1 if i is B.
 * d_split_energy (N x 4 doubles) receives half of fene / bexc / stack per particle
 * like the reference's split array. params: n_bond_types structs in host memory. */
int oxdna_bonded_forces(float* d_force, float* d_torque, float* d_virial, size_t virial_pitch,
                        double* d_split_energy, unsigned int N,
                        const double* d_pos, const double* d_orientation,
                        const oxdna_box* box,
                        const unsigned int* d_blist, const unsigned int* d_bpos,
                        unsigned int pitch, const unsigned int* d_n_bonds,
                        const oxdna_bonded_params* params, unsigned int n_bond_types,
                        void* stream);

/* ------------------------------------------------------------------------- */
/* (2) Context API: one system (or one batch of replicas) resident on one GPU.*/
/* ------------------------------------------------------------------------- */

/* Creates a context on HIP device `device`. stream may be NULL: the context then
 * creates its own (default-priority) stream. */
int oxdna_ctx_create(int device, void* stream, oxdna_ctx** out);
int oxdna_ctx_destroy(oxdna_ctx* ctx);

/* Particle state from host arrays: pos[3n], quat[4n] (w,x,y,z), type[n],
 * charge[n] (NULL = all 1.0, the reference default, oxdna/io.py:164-166). */
int oxdna_set_state(oxdna_ctx* ctx, unsigned int n, const double* pos, const double* quat,
                    const int32_t* type, const double* charge, const oxdna_box* box);
/* Velocities and angular momenta (body frame, quaternion form as in HOOMD:
 * angmom[4n]). NULL arrays are set to zero. */
int oxdna_set_momenta(oxdna_ctx* ctx, const double* vel, const double* angmom);
/* Bonds (A, B) with B the 5' neighbour of A (reference oxdna/io.py:218-220);
 * btype may be NULL (all type 0). Bonded pairs are excluded from every nonbonded
 * term (reference oxdna/model/dna1.py:307-309). */
int oxdna_set_bonds(oxdna_ctx* ctx, unsigned int nb, const int32_t* ab, const int32_t* btype);

/* Term parameters. term is one of oxdna_term except OXDNA_TERM_BONDED. The table is
 * ntypes*ntypes structs (row-major [ti][tj]), rcut is ntypes*ntypes COM cut-offs; any
 * field may differ between type pairs (the reference's param_type per type pair), each
 * type pair being evaluated with its own struct. Limits: ntypes is 1..4 (the oxDNA type
 * set A, G, C, T/U, which the bonded st_eps[4][4] table also assumes), and the distinct
 * structs beyond the first of every term together take at most 16 KiB (e.g. 40 distinct
 * fused HB + cross-stacking structs); beyond either the call fails with an error.
 * Passing params = NULL removes the term. */
int oxdna_set_pair_params(oxdna_ctx* ctx, int term, int ntypes, const void* params,
                          const double* rcut);
int oxdna_set_bond_params(oxdna_ctx* ctx, unsigned int n_bond_types,
                          const oxdna_bonded_params* params);

/* Neighbour list: cell list with skin `buffer` (reference hoomd.md.nlist.Cell(buffer,
 * exclusions=("bond",)), oxdna/model/dna1.py:307-309). r_list = max term cut-off +
 * buffer; rebuilt when any particle moved more than buffer/2 since the last build. */
int oxdna_set_nlist(oxdna_ctx* ctx, double buffer);

/* Evaluates the terms in term_mask (bits of oxdna_term) at the current state into
 * the context's force / torque / energy buffers. fused != 0 evaluates every
 * selected nonbonded term and the bonded term in one launch; fused == 0 launches
 * one kernel per term (the reference's structure). */
int oxdna_compute(oxdna_ctx* ctx, unsigned int term_mask, int fused);
/* Per-term total energies of the last compute (FP64, split_energy order). */
int oxdna_get_term_energies(oxdna_ctx* ctx, double out[8]);
/* Per-particle outputs of the last compute, FP64 host copies: force[4n]
 * (x, y, z, half energy), torque[3n], virial[6n] (xx, xy, xz, yy, yz, zz). */
int oxdna_get_forces(oxdna_ctx* ctx, double* force4, double* torque3, double* virial6);
/* The current neighbour list as sorted (i, j) pairs, i < j (host, 2*max_pairs ints);
 * *npairs receives the count. */
int oxdna_get_nlist_pairs(oxdna_ctx* ctx, int32_t* pairs, size_t max_pairs, size_t* npairs);

/* Langevin dynamics (translational + rotational, NO_SQUISH splitting). gamma and
 * gamma_r are per-particle-type drag coefficients (reference defaults 1.0). */
int oxdna_set_langevin(oxdna_ctx* ctx, double dt, double kT, double gamma, double gamma_r,
                       uint64_t seed);
/* Advances n Langevin steps (oxdna_set_langevin), each one fused force evaluation of
 * term_mask with the integrator in the same launch on systems up to ~2000 wavefronts
 * (a second kernel on larger ones). use_graph != 0 replays the steps 96 at a time as a
 * captured HIP graph; graph and direct launches give bit-identical trajectories. The
 * state (positions, orientations, momenta, forces) is complete when the call returns. */
int oxdna_run(oxdna_ctx* ctx, unsigned int nsteps, unsigned int term_mask, int use_graph);
int oxdna_get_state(oxdna_ctx* ctx, double* pos, double* quat, double* vel, double* angmom);
/* Batched replicas: the particles given to oxdna_set_state are n_replicas
 * independent systems. Replica r holds particles [offsets[r], offsets[r+1]) and lives
 * in its own periodic box boxes[3r..3r+2] (each length above twice the largest
 * cut-off); no pair or bond crosses replicas (bonds that would are rejected). Each
 * replica has at most 496 particles (its LDS layout fills one workgroup's 160 KiB). oxdna_compute / oxdna_run / oxdna_get_forces /
 * oxdna_get_state then act on all replicas at once, in one persistent launch per
 * chunk of steps (one workgroup per replica, its state resident in LDS); replica r
 * draws its Langevin noise with seed + r, i.e. exactly the stream of a one-system
 * context holding replica r alone with that seed. Term energies are summed over
 * the replicas; the virial is not computed (zero). n_replicas = 0 returns to one
 * system. A later oxdna_set_state with another particle count clears the replicas.
 * No counterpart in the reference (one system per HOOMD simulation there). */
int oxdna_set_replicas(oxdna_ctx* ctx, unsigned int n_replicas, const unsigned int* offsets,
                       const double* boxes);
/* Per-replica harmonic bias k[r] (|r_e - r_s| - d0[r])^2 on the centres of particles
 * s = ends[2r] and e = ends[2r+1] (indices local to replica r); k[r] = 0 leaves a
 * replica unbiased, NULL arrays remove the bias. Energy slot OXDNA_E_BIAS of
 * oxdna_get_replica_energies. Not a term of the reference: a new force for
 * umbrella-style ensembles. */
int oxdna_set_replica_bias(oxdna_ctx* ctx, const int32_t* ends, const double* k, const double* d0);
#define OXDNA_E_BIAS 8
#define OXDNA_N_REPLICA_ENERGY 9
/* Per-replica energies of the last compute: n_replicas x 9 doubles (the 8
 * split_energy slots, then the bias). */
int oxdna_get_replica_energies(oxdna_ctx* ctx, double* out);

/* Blocks until the context's stream is idle. */
int oxdna_synchronize(oxdna_ctx* ctx);
/* Measurement: times nrep launches of the fused force kernel for term_mask at the
 * current state with HIP events on the context's stream; *avg_ms = mean duration. */
int oxdna_time_forces(oxdna_ctx* ctx, unsigned int term_mask, unsigned int nrep, double* avg_ms);
/* Measurement: advances nsteps Langevin steps like oxdna_run with use_graph = 1
 * (nsteps >= 97) and returns in *avg_ms the mean duration of one graph-launched inner
 * step, timed with HIP events on the context's stream around the graph replays: the
 * fused force + integrator launch a run step is (on systems too large for that
 * fusion, the force launch and the integrator kernel). No counterpart in the
 * reference (measurement of this engine's run step). */
int oxdna_time_steps(oxdna_ctx* ctx, unsigned int term_mask, unsigned int nsteps, double* avg_ms);
/* Measurement: directed evaluation counts of the current state, out[2*t] = pairs of
 * term t inside its cut-off, out[2*t+1] = those with nonzero energy (t = bonded:
 * bond endpoints, nonzero = stacking energy != 0). */
int oxdna_count_evaluations(oxdna_ctx* ctx, unsigned int term_mask, uint64_t out[14]);
/* Statistics: number of neighbour-list builds and the current list capacity. */
int oxdna_get_stats(oxdna_ctx* ctx, uint64_t* n_builds, uint32_t* max_neigh, uint64_t* step);
/* Lanes per particle of the force kernel: 4, 8, 16 or 32 (a workgroup of 256 threads
 * then owns 64, 32, 16 or 8 particles), or 0 (default) to let the autotuner time 4, 8
 * and 16 at the first run. The choice sets the order of each particle's force sums, so
 * reruns are bit-identical only with the same value: pin it for bit-reproducible
 * trajectories across processes. */
int oxdna_set_lanes_per_particle(oxdna_ctx* ctx, int lanes);
/* Replica kernel shape, pinned: workgroup of 64, 128, 256 or 512 threads (at least the
 * largest replica's size rounded up to a power of two, checked at the next run) and a
 * tile depth of the LDS share's pair-record capacity / depth_divisor (1, 2 or 4; at
 * least the largest replica + 64 records): a shallower tile runs more passes over
 * ranges of first ends per step and leaves LDS for more workgroups per compute unit.
 * workgroup 0 returns both to the autotuner. Replica sets with per-type-pair parameter
 * tables, and replicas above 488 particles, always run 512-thread workgroups at full
 * depth. Plays the part of setting
 * HOOMD's AutotunedObject.kernel_parameters, the interface of the Autotuner the
 * reference registers its kernels with (src/OxDNABondedForceComputeGPU.cc:35-37). */
int oxdna_set_replica_tiling(oxdna_ctx* ctx, int workgroup, int depth_divisor);
/* Autotuner choices: lanes per particle of the force kernel (4, 8, 16 or 32) and the
 * workgroup size of the replica kernel (64..512; 0 before the first replica run). */
int oxdna_get_tuning(oxdna_ctx* ctx, int* lanes_per_particle, int* replica_workgroup);

/* Every autotuned launch parameter of a context (oxdna_get_tuning_ex). */
typedef struct oxdna_tuning {
    int lanes_per_particle;       /* force kernel: lanes per particle (4, 8, 16, 32) */
    int particles_per_workgroup;  /* force kernel: particles per 256-thread workgroup (its tile) */
    int replica_workgroup;        /* replica kernel: threads per workgroup (0 before the first replica run) */
    int replica_records_per_pass; /* replica kernel: pair records per pass, its tile depth (0: none yet) */
    double nlist_buffer;          /* neighbour-list skin in use (oxdna_set_nlist or the skin autotuner) */
} oxdna_tuning;
/* The autotuner's current choices. Block size and tile depth are timed at the first run
 * of a setup (the force kernel's lanes per particle; the replica kernel's workgroup size
 * and records per pass); the skin by oxdna_tune_nlist_buffer. The reference tunes its
 * kernels' block sizes with HOOMD's Autotuner (src/OxDNABondedForceComputeGPU.cc:35-37,
 * 91-100) and fixes the skin (buffer 0.4, oxdna/model/dna2.py:207-208). */
int oxdna_get_tuning_ex(oxdna_ctx* ctx, oxdna_tuning* out);
/* Neighbour-list skin autotuner (one-system contexts with an integrator and at least
 * one nonbonded term in term_mask). From the current state, runs min(steps / 4, 400)
 * untimed and then `steps` timed Langevin steps at each of the n_buffers candidate
 * buffers (each in (0, 4]; buffers = NULL: 0.1, 0.2, ..., 0.8, n_buffers ignored),
 * rows and capacities built afresh for each, and keeps the one with the most steps/s
 * (host clock around whole runs: rebuilds and chunk ends included). steps_per_s
 * (NULL or n_buffers doubles, 8 for the default set) receives each candidate's rate,
 * *chosen (may be NULL) the kept buffer, which stays set as by oxdna_set_nlist. The
 * state and step count are restored afterwards: the trajectory continues from the
 * state it had, at the chosen skin. Call it on an equilibrated state: the rebuild rate
 * it measures is that state's. */
int oxdna_tune_nlist_buffer(oxdna_ctx* ctx, unsigned int term_mask, const double* buffers,
                            unsigned int n_buffers, unsigned int steps, double* steps_per_s,
                            double* chosen);

#ifdef __cplusplus
}
#endif
#endif /* OXDNA_HIP_H */
