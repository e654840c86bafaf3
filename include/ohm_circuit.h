/*
 * ohm_circuit.h — plain-C data formats shared by the MI355X solve backend and
 * its CPU oracle: the flat circuit encoding fed to the MNA compiler, the
 * compiled MNA system, and the per-device descriptors.
 *
 * The descriptor structs are byte-for-byte the reference's `repr(C)` Pod
 * layouts (src/compiler.rs:52-149; sizes 48/96/64 B asserted at
 * src/compiler.rs:1545,1575,1584) so a caller can hand the reference
 * compiler's descriptor slices straight to this library.
 */
#ifndef OHM_CIRCUIT_H
#define OHM_CIRCUIT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: one per OhmnivoreError variant (src/error.rs:4-36), plus the
 * two conditions a C ABI needs that Rust expresses through types. */
typedef enum {
  OHM_OK = 0,
  OHM_ERR_PARSE = 1,
  OHM_ERR_COMPILE = 2,
  OHM_ERR_SOLVE = 3,
  OHM_ERR_NEWTON_NOT_CONVERGED = 4,
  OHM_ERR_NEWTON_NUMERICAL = 5,
  OHM_ERR_ANALYSIS = 6,
  OHM_ERR_CONVERGENCE = 7,
  OHM_ERR_TIMESTEP = 8,
  OHM_ERR_IO = 9,
  OHM_ERR_NO_DEVICE = 10, /* "no GPU adapter" (tests/diode_integration_test.rs:149) */
  OHM_ERR_INVALID = 11    /* bad handle / argument (a Rust type error) */
} ohm_status;

#define OHM_SENTINEL 0xFFFFFFFFu /* u32::MAX ground sentinel (src/compiler.rs:601) */

/* Component kinds, one per ir::Component variant (src/ir.rs:11-62). */
enum {
  OHM_COMP_R = 0,
  OHM_COMP_C = 1,
  OHM_COMP_L = 2,
  OHM_COMP_V = 3,
  OHM_COMP_I = 4,
  OHM_COMP_D = 5,
  OHM_COMP_Q = 6,
  OHM_COMP_M = 7
};

#define OHM_COMP_HAS_DC 1
#define OHM_COMP_HAS_AC 2

/* One netlist component. Node references index the circuit's node-name
 * table; ground is any name equal to "0" or case-insensitively "GND"
 * (src/compiler.rs:178-180). `model` indexes the model table of the
 * component's kind, or is -1 when the named model is undefined. */
typedef struct {
  int32_t kind;
  int32_t node[3]; /* two-terminal parts use node[0], node[1] */
  int32_t model;
  int32_t flags;   /* OHM_COMP_HAS_DC | OHM_COMP_HAS_AC for V/I sources */
  double value;    /* R, C, L value; V/I DC value */
  double ac_mag;
  double ac_phase_deg;
} ohm_component;

typedef struct { double is, n; } ohm_diode_model;                                 /* src/ir.rs:67-74 */
typedef struct { double is, bf, br, nf, nr; int32_t is_npn, _pad; } ohm_bjt_model; /* src/ir.rs:76-92 */
typedef struct { double vto, kp, lambda; int32_t is_nmos, _pad; } ohm_mosfet_model; /* src/ir.rs:94-104 */

typedef struct {
  int32_t n_components;
  const ohm_component *components;
  const char *const *component_names; /* used in error messages */
  const char *const *model_names;     /* per component: referenced model name or NULL */
  int32_t n_node_names;
  const char *const *node_names;
  int32_t n_diode_models, n_bjt_models, n_mosfet_models;
  const ohm_diode_model *diode_models;
  const ohm_bjt_model *bjt_models;
  const ohm_mosfet_model *mosfet_models;
} ohm_circuit;

/* GpuDiodeDescriptor (src/compiler.rs:52-73): 48 bytes. */
typedef struct {
  uint32_t anode_idx, cathode_idx;
  float is_val, n_vt;
  uint32_t g_row_col[4]; /* [aa, cc, ac, ca] */
  uint32_t b_idx[2];
  uint32_t _padding[2];
} ohm_diode_desc;

/* GpuBjtDescriptor (src/compiler.rs:80-113): 96 bytes. */
typedef struct {
  uint32_t collector_idx, base_idx, emitter_idx;
  float polarity, is_val, bf, br, nf_vt, nr_vt;
  uint32_t node_fixed[3];
  uint32_t g_row_col[9]; /* [CC, CB, CE, BC, BB, BE, EC, EB, EE] */
  uint32_t b_idx[3];
} ohm_bjt_desc;

/* GpuMosfetDescriptor (src/compiler.rs:120-149): 64 bytes. */
typedef struct {
  uint32_t drain_idx, gate_idx, source_idx;
  float polarity, vto, kp, lambda;
  uint32_t node_fixed_packed; /* bit0 drain, bit1 gate, bit2 source */
  uint32_t g_row_col[6];      /* [DD, DG, DS, SD, SG, SS] */
  uint32_t b_idx[2];
} ohm_mosfet_desc;

/* MnaSystem (src/compiler.rs:153-175). Arrays are owned by whoever filled
 * the struct and released with that library's matching free call. CSR
 * indices are u32 as in the reference's GPU upload (src/solver/gpu.rs:124-125). */
typedef struct {
  int64_t size, n_nodes, n_branches;
  int64_t g_nnz, c_nnz;
  uint32_t *g_rowptr, *g_col;
  double *g_val;
  uint32_t *c_rowptr, *c_col;
  double *c_val;
  double *b_dc;
  double *b_ac;             /* interleaved (re, im), length 2*size */
  int32_t *node_order;      /* matrix index -> node-name table index */
  int32_t *branch_component;/* branch k -> component index */
  int64_t n_diodes, n_bjts, n_mosfets;
  ohm_diode_desc *diodes;
  ohm_bjt_desc *bjts;
  ohm_mosfet_desc *mosfets;
} ohm_mna;

#ifdef __cplusplus
}
#endif
#endif /* OHM_CIRCUIT_H */
