/*
 * geofdtd.h -- C ABI of the MI355X (gfx950) GeodesicFDTD time-step library.
 *
 * The library advances the four staggered field families of a geodesic
 * Earth-ionosphere FDTD model (Er on vertices x radial nodes, Ht on edges x
 * nodes, Et on edges x radial midpoints, Hr on faces x midpoints) with two
 * fused HIP kernels per leapfrog step.  Everything crossing this boundary is
 * a plain pointer, size or scalar: host tables are read once by
 * geo_plan_create (FP64 / int64, as the reference builds them), field state
 * lives in caller-owned device buffers bound with geo_plan_bind.
 *
 * Reference interfaces replaced (paths relative to the ruddyscent/
 * ionosphere-fdtd source tree, src/ionosphere_fdtd/):
 *   geo_plan_create   <- GeodesicFDTD.__init__ table setup: solver.py:315-375,
 *                        _prepare_geometry solver.py:408-442,
 *                        _prepare_material_coefficients solver.py:706-761,
 *                        NumPyBackend._vertex_incidence numpy_backend.py:46-76,
 *                        SurfaceImpedanceADE.__init__ surface_impedance.py:176-203,
 *                        GeodesicPlasmaCoupler.__init__ plasma.py:458-487,
 *                        source staggered_distribution sources.py:220-235
 *   geo_plan_bind     <- the solver's field attributes er/et/hr/ht
 *                        (solver.py:315-322), SI memory (surface_impedance.py:203),
 *                        plasma current_density (plasma.py:411); re-callable
 *                        like checkpoint.py:259-312 replacing them
 *   geo_step          <- GeodesicFDTD.step (solver.py:789-816) and the compiled
 *                        chunk path backend.compile_step (solver.py:365-375,
 *                        torch_backend.py:91-96): _advance_fields
 *                        (solver.py:867-989) for n steps, full chunks replayed
 *                        as one HIP graph, remainder as single steps
 *   geo_step_prepare  <- the compile half of backend.compile_step
 *                        (solver.py:365-375): the chunk graph is built
 *                        before the first geo_step call that replays it
 *   geo_sample_er     <- record_er_observations' sample() (solver.py:1122-1124)
 *   geo_set_receivers <- record_er_observations argument upload
 *                        (solver.py:1097-1120)
 *   geo_diagnostics   <- diagnostics() max_abs (solver.py:1027-1030) and the
 *                        FP64 energy / conductive-loss reductions of
 *                        verification/physics_diagnostics/model.py:242-267
 *   geo_radial_profiles <- the per-layer energy and loss profiles of the same
 *                        sampler (model.py:136-186, 242-267)
 *   geo_last_error    <- the Python exceptions the reference raises
 *
 * Error convention: 0 = success; GEO_ERR_INVALID maps to ValueError,
 * GEO_ERR_UNAVAILABLE to BackendUnavailableError (reference base.py:15-16),
 * GEO_ERR_HIP to RuntimeError.  geo_last_error() returns the message of the
 * calling thread's last failure.
 *
 * Threading: one host thread per plan.  All device work is enqueued on the
 * caller's stream; no call synchronises the host except geo_diagnostics,
 * geo_radial_profiles, building a step graph (geo_step_prepare, or the
 * first geo_step of a chunk size), which waits for the graph's upload to the
 * device, and geo_set_receivers / geo_set_h_receivers, which replace their
 * device tables with hipFree, hipMalloc and blocking copies (call them
 * between steps, not while work that reads the old tables is queued on
 * another stream).
 */
#ifndef GEOFDTD_H
#define GEOFDTD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GEO_OK 0
#define GEO_ERR_INVALID (-1)
#define GEO_ERR_UNAVAILABLE (-2)
#define GEO_ERR_HIP (-3)

/* Field storage type.  Update arithmetic is FP32 for F32/F16/BF16 and FP64
 * for F64. */
#define GEO_STORAGE_F32 0
#define GEO_STORAGE_F64 1
#define GEO_STORAGE_F16 2
#define GEO_STORAGE_BF16 3

#define GEO_MAPPING_AUTO 0
#define GEO_MAPPING_ROWS 1
#define GEO_MAPPING_FLAT 2

#define GEO_MAX_SOURCE 16
#define GEO_MAX_SI_TERMS 32
#define GEO_MAX_SPECIES 4

/* Topology and unit-sphere metric of the geodesic mesh (mesh.py:22-41). */
typedef struct GeoMeshDesc {
  int64_t n_vertices;
  int64_t n_edges;
  int64_t n_faces;
  const int64_t* edges;            /* (E,2) tail, head */
  const int64_t* faces;            /* (F,3) */
  const int64_t* face_edges;       /* (F,3) */
  const int64_t* face_edge_signs;  /* (F,3) +1 / -1 */
  const int64_t* edge_left_faces;  /* (E,) */
  const int64_t* edge_right_faces; /* (E,) */
  const int64_t* vertex_edges;     /* (V,max_degree) padded incidence */
  const int64_t* vertex_edge_signs;/* (V,max_degree) +1 tail, -1 head, 0 pad */
  int64_t max_degree;              /* <= 6 */
  const double* primal_edge_angles;     /* (E,) theta */
  const double* dual_edge_angles;       /* (E,) phi */
  const double* face_solid_angles;      /* (F,) Omega */
  const double* dual_cell_solid_angles; /* (V,) A */
} GeoMeshDesc;

/* Radial grid (solver.py:260-283). */
typedef struct GeoRadialDesc {
  int64_t radial_cells;      /* R */
  const double* radii_m;     /* (R+1,) */
  int32_t full_spherical;    /* 1 full-spherical, 0 thin-shell */
  int32_t reserved;
} GeoRadialDesc;

/* Time step, loss coefficients, boundary, sources and plasma. */
typedef struct GeoPhysicsDesc {
  int32_t storage;           /* GEO_STORAGE_* */
  int32_t scaling;           /* narrow storage: 0 static scales, 1 dynamic power-of-two
                                rescaling before every chunk of steps */
  double time_step_s;
  /* Exponential / trapezoidal loss coefficients.  coef_rows = 1 means one
   * horizontally uniform radial profile (solver.py:737-748). */
  const double* ca_er; const double* cb_er; int64_t er_coef_rows; /* (rows,R+1) */
  const double* ca_et; const double* cb_et; int64_t et_coef_rows; /* (rows,R) */
  /* Absolute permittivity (F/m) with the same rows as the coefficients, for
   * the energy diagnostic; NULL means vacuum. */
  const double* eps_er; const double* eps_et;
  /* Per-field storage scale factors (powers of two; 1 for F32/F64). */
  double scale_er, scale_et, scale_hr, scale_ht;
  /* Surface-impedance lower boundary; si_terms = 0 selects PEC. */
  int32_t si_terms;
  int32_t reserved1;
  const double* si_decay;            /* (terms,) (1-a)/(1+a) */
  const double* si_drive;            /* (terms,) a/(1+a) */
  const double* si_history_weights;  /* (terms,) w/(1+a) */
  const double* si_scale;            /* (si_scale_rows,) sqrt(mu/sigma) */
  int64_t si_scale_rows;             /* 1 or E */
  /* Vertical source support (sources.py:220-235). */
  int32_t n_source;
  int32_t reserved2;
  const int64_t* source_vertices;    /* (n,) */
  const int64_t* source_layers;      /* (n,) radial nodes */
  const double* source_weights;      /* (n,) */
  double source_length_m;            /* vertical_element_length_m */
  /* Tangential source support (sources.py:290-335). */
  int32_t n_tsource;
  int32_t reserved3;
  const int64_t* tsource_edges;
  const int64_t* tsource_layers;     /* radial midpoints */
  const double* tsource_weights;
  /* Cold-plasma ADE (plasma.py:371-527).  n_species = 0 disables it. */
  int32_t n_species;
  int32_t magnetized;                /* 0: B == 0 everywhere (Drude form) */
  int32_t band_lo, band_hi;          /* midpoint layers with plasma state */
  const double* plasma_reconstruction; /* (F,3,3) pinv(tangents).T */
  const double* face_tangents;       /* (F,3,3) unit edge tangents of each face */
  const double* face_centers;        /* (F,3) */
  const double* left_tangents;       /* (E,3) */
  const double* right_tangents;      /* (E,3) */
  const int64_t* vertex_faces;       /* (V,vf_width) */
  const double* vertex_face_weights; /* (V,vf_width) */
  int64_t vf_width;                  /* <= 6 */
  /* ADE coefficients decay, cos, sin, d_parallel, d_real, d_imag for every
   * species: (n_species, 6, plasma_coef_rows, R); rows = 1 or F. */
  const double* plasma_coef;
  int64_t plasma_coef_rows;
  const double* plasma_b;            /* (plasma_b_rows, R, 3) unit B */
  int64_t plasma_b_rows;             /* 1 or F */
  /* Conductivity (S/m) with the same rows as the coefficients, for the
   * conductive-loss diagnostic; NULL means no loss. */
  const double* sigma_er; const double* sigma_et;
  /* Narrow storage rounding of every stored value: 0 to nearest; 1 and 2
   * stochastically (unbiased), with the threshold of each element hashed
   * from (step, field, element) independently every step (1) or a hashed
   * per-element phase advanced by the golden ratio each step (2, a
   * low-discrepancy sequence).  Runs are repeatable bit for bit.  Ignored
   * for F32/F64. */
  int32_t storage_rounding;
  /* Thread mapping of the two step kernels (GEO_MAPPING_*): the strips of
   * each block either belong to whole rows (an integer number of rows per
   * 256-thread block) or run through the rows flat (rows may straddle
   * blocks, every lane busy).  Both give the same bits; AUTO takes FLAT for
   * a kernel whose whole-row blocks would leave at least 16 lanes idle in
   * one of its segments. */
  int32_t thread_mapping;
} GeoPhysicsDesc;

/* Device buffers owned by the caller.  Pitches count elements. */
typedef struct GeoState {
  void* er; int64_t er_pitch;        /* (V, er_pitch), storage type */
  void* ht; int64_t ht_pitch;        /* (E, ht_pitch) */
  void* et; int64_t et_pitch;        /* (E, et_pitch) */
  void* hr; int64_t hr_pitch;        /* (F, hr_pitch) */
  void* si_memory;                   /* (E, si_pitch) edge-major, compute type; columns past
                                        si_terms are zero */
  void* plasma_current;              /* (n_species, components, F, plasma_pitch), compute */
  int64_t plasma_pitch;
} GeoState;

/* Layout a caller must allocate for a plan (all counts in elements). */
typedef struct GeoLayout {
  int64_t node_pitch;      /* pitch for Er / Ht rows (>= R+1) */
  int64_t mid_pitch;       /* pitch for Et / Hr rows (>= R) */
  int64_t band_lo;         /* first plasma midpoint layer stored */
  int64_t band_hi;         /* one past the last */
  int64_t plasma_pitch;    /* pitch of plasma rows (>= band_hi - band_lo) */
  int64_t plasma_components; /* state values per face, layer and species (3) */
  int64_t vector_elems;    /* elements per 16-byte strip */
  int64_t compute_bytes;   /* 4 (FP32 arithmetic) or 8 (FP64) */
  int64_t si_pitch;        /* row pitch of the surface-impedance memory (16-byte rows) */
  int64_t plasma_basis;    /* 1 = (J.n, J.t0, J.t1) per face, n the face centre and t_c
                              its edge tangents (B = 0); 0 = Cartesian J (magnetised) */
  int64_t radial_closed_form; /* 1 = uniform radii and FP32 arithmetic: the kernels form
                                 1/r and the radial-derivative weights from r_0 and dr
                                 instead of reading per-layer tables */
  int64_t h_thread_mapping; /* GEO_MAPPING_ROWS or GEO_MAPPING_FLAT taken by the H kernel */
  int64_t e_thread_mapping; /* ... and by the E kernel */
} GeoLayout;

typedef struct GeoDiagOut {
  double max_abs_er, max_abs_et, max_abs_hr, max_abs_ht;
  double energy_er, energy_et, energy_hr, energy_ht; /* 0.5*eps*E^2*w, 0.5*mu0*H^2*w */
  int64_t nonfinite;       /* count of non-finite field values */
  /* Discrete div-B of the prism cells (face f, node k):
   * D = Hr(f,k) Omega rbar_k^2 - Hr(f,k-1) Omega rbar_{k-1}^2
   *     - sum_c s_fc theta_c r_k L_k Ht(e_c,k), with Hr(f,-1) = Hr(f,R) = 0;
   * cells at k = 0 are skipped under a surface-impedance ground. */
  double divb_max_abs;     /* max |D| */
  double divb_flux_scale;  /* max over cells of the summed |flux| terms */
  int64_t storage_overflow; /* 1 once a rescale point saw a non-finite stored value */
  double loss_er, loss_et;  /* conductive loss sigma*E^2*w (W) */
  double sumsq_er, sumsq_et, sumsq_hr, sumsq_ht; /* sum of squares of the finite
                                                     (entity, layer) values, for RMS */
  int64_t nonfinite_er, nonfinite_et, nonfinite_hr, nonfinite_ht; /* per field */
} GeoDiagOut;

/* Host arrays for geo_radial_profiles; a NULL pointer skips that profile.
 * Lengths: er / ht R+1 nodes, et / hr R midpoints. */
typedef struct GeoProfiles {
  double* energy_er; double* energy_et; double* energy_hr; double* energy_ht; /* J per layer */
  double* loss_er; double* loss_et;                                           /* W per layer */
} GeoProfiles;

typedef struct GeoPlan GeoPlan;

int geo_plan_create(const GeoMeshDesc* mesh, const GeoRadialDesc* radial,
                    const GeoPhysicsDesc* physics, GeoPlan** out);
int geo_plan_layout(const GeoPlan* plan, GeoLayout* out);
int geo_plan_bind(GeoPlan* plan, const GeoState* state);
/* Advance n_steps leapfrog steps.  currents: device array of n_steps source
 * currents (compute type) evaluated at (n+1/2)dt, or NULL for zero.  chunk > 1
 * replays full chunks as a HIP graph.  trace: optional device buffer of
 * (n_steps+1, n_receivers) compute-type rows; row 0 is the state before the
 * first step (record_er_observations semantics). */
int geo_step(GeoPlan* plan, const void* currents, int64_t n_steps, int32_t chunk,
             void* trace, void* stream);
/* Capture and instantiate the HIP graph of `chunk` steps that geo_step
 * replays for that chunk, without running it (the compile half of
 * backend.compile_step, solver.py:365-375, done ahead of the first call).
 * The graph stays valid until geo_plan_bind or geo_set_receivers changes the
 * kernel parameters. */
int geo_step_prepare(GeoPlan* plan, int32_t chunk);
/* Step index of the bound state (GeodesicFDTD.steps; the plan counts the
 * steps it advances itself) for the stochastic rounding of narrow storage:
 * a state restored from a checkpoint at step n rounds as the run that was
 * never interrupted.  No reference counterpart (the reference stores FP32 /
 * FP64 only).  Enqueued on `stream`. */
int geo_plan_set_step_index(GeoPlan* plan, int64_t step_index, void* stream);
/* Receivers for trace recording: (n,3) vertices, (n,) node layers, (n,3)
 * weights. */
int geo_set_receivers(GeoPlan* plan, const int64_t* vertices, const int64_t* layers,
                      const double* weights, int32_t n_receivers);
/* H receivers for geo_step_h: (n_face_receivers, face_width) faces, radial
 * midpoint layers and weights, and (n_edge_receivers, edge_width) edges,
 * radial node layers and weights.  Replaces the argument validation and
 * upload of record_h_observations (solver.py:1158-1203). */
int geo_set_h_receivers(GeoPlan* plan, const int64_t* faces, const int64_t* face_layers,
                        const double* face_weights, int32_t n_face_receivers, int32_t face_width,
                        const int64_t* edges, const int64_t* edge_layers,
                        const double* edge_weights, int32_t n_edge_receivers, int32_t edge_width);
/* Advance n_steps like geo_step while writing H traces
 * (record_h_observations, solver.py:1138-1229): hr_trace (rows, n_face_receivers)
 * and ht_trace (rows, n_edge_receivers), compute type, either may be NULL,
 * with rows = n_steps / sample_every + 1 (+1 when sample_every does not
 * divide n_steps).  Row 0 is the H state before the first step, row k the
 * state after step k * sample_every, the last row the state after the last
 * step: H**(k-1/2) in the reference's half-step labelling.  Each sample is
 * sum_q w[i,q] * H[entity[i,q], layer[i,q]], written by one block of the E
 * launch (which does not write H), so graph chunks replay unchanged. */
int geo_step_h(GeoPlan* plan, const void* currents, int64_t n_steps, int32_t chunk,
               int32_t sample_every, void* hr_trace, void* ht_trace, void* stream);
/* Write one trace row (n_receivers compute-type values) from the current Er. */
int geo_sample_er(GeoPlan* plan, void* row, void* stream);
/* FP64 reductions over the bound state; synchronises the stream. */
int geo_diagnostics(GeoPlan* plan, GeoDiagOut* out, void* stream);
/* Per-layer FP64 energy and conductive-loss profiles (the horizontal sums of
 * verification/physics_diagnostics/model.py:242-267 times the radial
 * weights r^2 L_k at nodes, rbar^2 dr at midpoints); synchronises the
 * stream. */
int geo_radial_profiles(GeoPlan* plan, const GeoProfiles* out, void* stream);
/* Average kernel durations (ms) of n_steps individually launched and
 * event-timed steps: out_ms[0] = H kernel, out_ms[1] = E kernel. */
int geo_time_kernels(GeoPlan* plan, int64_t n_steps, void* stream, double* out_ms);
/* Current storage scales (stored = value * scale) of er, et, hr, ht; 1 for
 * FP32/FP64 storage.  Synchronises the stream.  No reference counterpart: the
 * reference stores fields in FP32/FP64 only. */
int geo_get_storage_scales(GeoPlan* plan, double* out4, void* stream);
/* Replace the storage scales (powers of two in [2^-60, 2^60]) after the
 * caller has written field values at those scales.  Synchronises the stream. */
int geo_set_storage_scales(GeoPlan* plan, const double* in4, void* stream);
void geo_plan_destroy(GeoPlan* plan);
/* Bandwidth probe (no counterpart in the reference): average time (ms) of a
 * kernel moving one 16-byte strip per thread over `bytes` -- mode 0 copy src->dst, 1 read src,
 * 2 write dst, +4 with non-temporal hints; mode 8 is the access pattern of
 * the H kernel's edge strips without arithmetic, in the step kernels' XCD block
 * order (rows of 68 floats in dst
 * updated in place from two gathered rows of src, a third as many rows);
 * mode 9 is the same with the two src rows read from an int32 index table
 * that the probe allocates and fills, so that each gather waits on an index
 * load as the step kernels' gathers do.
 * The measured rates are the achievable HBM ceilings reported beside the
 * roofline, and the copy's rocprofv3 FETCH_SIZE / WRITE_SIZE readings
 * calibrate the counters. */
int geo_stream_probe(void* dst, const void* src, int64_t bytes, int32_t mode, int32_t repeats,
                     void* stream, double* out_ms);
/* Ensembles (no reference counterpart: the reference advances one
 * GeodesicFDTD per process, solver.py:789-816).  Independent runs on the
 * same mesh -- each a bound plan with its own radial grid, time step,
 * coefficients, plasma, ground, source and receivers -- advance together:
 * one H and one E launch per step cover every run, full chunks replay as one
 * HIP graph.  Plans must outlive the ensemble and share the storage type and
 * plasma kind; dynamically scaled FP16 plans are not accepted.  Results are
 * bitwise those of stepping every plan alone. */
typedef struct GeoEnsemble GeoEnsemble;
int geo_ensemble_create(GeoPlan* const* plans, int32_t n_plans, GeoEnsemble** out);
/* Column-block shape: strips_per_thread 1, 2 or 4 (16-byte strips a thread
 * updates per row); block_order bit 0: 0 run-major, 1 runs interleaved;
 * bit 1 (strips_per_thread 1 only): every run takes the flat thread mapping
 * (GEO_MAPPING_FLAT) that its flat indices allow (rows x strips < 2^31 per
 * kernel), where otherwise each run keeps its plan's mapping.
 * All shapes give the same bits. */
int geo_ensemble_configure(GeoEnsemble* ens, int32_t strips_per_thread, int32_t block_order);
/* Autotuner: time every shape over `steps` source-free steps, keep the
 * fastest and restore every run's state.  Synchronises the stream. */
int geo_ensemble_tune(GeoEnsemble* ens, int32_t steps, void* stream, int32_t* strips_per_thread,
                      int32_t* block_order, double* ms_per_step);
/* Advance every run n_steps.  currents: (n_plans, n_steps) compute type or
 * NULL; traces: NULL or (n_plans, n_steps + 1, n_receivers) with the same
 * receiver count in every run (geo_set_receivers on each plan). */
int geo_ensemble_step(GeoEnsemble* ens, const void* currents, int64_t n_steps, int32_t chunk,
                      void* traces, void* stream);
void geo_ensemble_destroy(GeoEnsemble* ens);
/* Surface operators of the backend interface (backends/base.py:38-56, with
 * the semantics and summation order of numpy_backend.py:90-114) on device
 * arrays with `cols` contiguous values per entity, FP32 (value_bytes 4) or
 * FP64 (8).  The operator context keeps the mesh topology on the device.
 *   GEO_OP_EDGE_DIFFERENCE       (V,cols) -> (E,cols)  v[head] - v[tail]
 *   GEO_OP_DUAL_EDGE_DIFFERENCE  (F,cols) -> (E,cols)  f[left] - f[right]
 *   GEO_OP_FACE_CIRCULATION      (E,cols) -> (F,cols)  sum_c s_fc e[face_edges]
 *   GEO_OP_DUAL_CELL_CIRCULATION (E,cols) -> (V,cols)  sum_slot s e[vertex_edges]
 * geo_array_diff is backend.diff of a (rows, cols) array along axis 0 or 1. */
#define GEO_OP_EDGE_DIFFERENCE 0
#define GEO_OP_DUAL_EDGE_DIFFERENCE 1
#define GEO_OP_FACE_CIRCULATION 2
#define GEO_OP_DUAL_CELL_CIRCULATION 3
typedef struct GeoMeshOps GeoMeshOps;
int geo_mesh_ops_create(const GeoMeshDesc* mesh, GeoMeshOps** out);
int geo_mesh_op(GeoMeshOps* ops, int32_t op, int32_t value_bytes, const void* in, int64_t cols,
                void* out, void* stream);
int geo_array_diff(int32_t value_bytes, const void* in, int64_t rows, int64_t cols, int32_t axis,
                   void* out, void* stream);
void geo_mesh_ops_destroy(GeoMeshOps* ops);
const char* geo_last_error(void);
int geo_version(void);

#ifdef __cplusplus
}
#endif

#endif /* GEOFDTD_H */
