"""Circuit IR and seeded synthetic circuit generators.

The IR mirrors the reference's ``ir::Circuit`` (src/ir.rs:11-178): components
in netlist order, named nodes (ground is "0" or case-insensitive "GND"), and
named diode/BJT/MOSFET models looked up by first name match
(src/compiler.rs:416-424). Components are kept in a numpy structured array
whose layout is the C ``ohm_component`` of include/ohm_circuit.h, so the
compiler receives the netlist without per-element marshalling.

The generators build the five benchmark circuits. Their random draws come
from one SplitMix64 stream per circuit family (seed = family number),
consumed in netlist emission order, so component order fixes node order and
therefore the CSR pattern.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

# Component kinds (include/ohm_circuit.h; one per src/ir.rs:11-62 variant).
R, C, L, V, I, D, Q, M = range(8)
HAS_DC, HAS_AC = 1, 2

COMPONENT_DTYPE = np.dtype(
    [
        ("kind", "<i4"),
        ("node", "<i4", (3,)),
        ("model", "<i4"),
        ("flags", "<i4"),
        ("value", "<f8"),
        ("ac_mag", "<f8"),
        ("ac_phase", "<f8"),
    ],
    align=True,
)
assert COMPONENT_DTYPE.itemsize == 48

DIODE_MODEL_DTYPE = np.dtype([("is", "<f8"), ("n", "<f8")], align=True)
BJT_MODEL_DTYPE = np.dtype(
    [("is", "<f8"), ("bf", "<f8"), ("br", "<f8"), ("nf", "<f8"), ("nr", "<f8"),
     ("is_npn", "<i4"), ("_pad", "<i4")], align=True)
MOSFET_MODEL_DTYPE = np.dtype(
    [("vto", "<f8"), ("kp", "<f8"), ("lambda", "<f8"), ("is_nmos", "<i4"), ("_pad", "<i4")],
    align=True)

# Descriptor layouts: byte-for-byte src/compiler.rs:52-149.
DIODE_DESC_DTYPE = np.dtype(
    [("anode_idx", "<u4"), ("cathode_idx", "<u4"), ("is_val", "<f4"), ("n_vt", "<f4"),
     ("g_row_col", "<u4", (4,)), ("b_idx", "<u4", (2,)), ("_padding", "<u4", (2,))])
BJT_DESC_DTYPE = np.dtype(
    [("collector_idx", "<u4"), ("base_idx", "<u4"), ("emitter_idx", "<u4"),
     ("polarity", "<f4"), ("is_val", "<f4"), ("bf", "<f4"), ("br", "<f4"), ("nf_vt", "<f4"),
     ("nr_vt", "<f4"), ("node_fixed", "<u4", (3,)), ("g_row_col", "<u4", (9,)),
     ("b_idx", "<u4", (3,))])
MOSFET_DESC_DTYPE = np.dtype(
    [("drain_idx", "<u4"), ("gate_idx", "<u4"), ("source_idx", "<u4"), ("polarity", "<f4"),
     ("vto", "<f4"), ("kp", "<f4"), ("lambda", "<f4"), ("node_fixed_packed", "<u4"),
     ("g_row_col", "<u4", (6,)), ("b_idx", "<u4", (2,))])
assert DIODE_DESC_DTYPE.itemsize == 48
assert BJT_DESC_DTYPE.itemsize == 96
assert MOSFET_DESC_DTYPE.itemsize == 64

SENTINEL = 0xFFFFFFFF

# ohm_tran_source (include/ohm_solver.h): SourceEntry, src/analysis/transient.rs:200-208
TRAN_SOURCE_DTYPE = np.dtype(
    [("kind", "<i4"), ("n_idx", "<i4"), ("idx", "<u4", (2,)), ("sign", "<f8", (2,)),
     ("dc", "<f8"), ("param", "<f8", (7,)), ("pwl_offset", "<i8"), ("pwl_count", "<i8")])
assert TRAN_SOURCE_DTYPE.itemsize == 112
SRC_DC, SRC_PULSE, SRC_SIN, SRC_EXP, SRC_PWL = range(5)


@dataclass
class Pulse:
    """PULSE(v1 v2 td tr tf pw per) — src/ir.rs:108-116."""
    v1: float
    v2: float
    td: float
    tr: float
    tf: float
    pw: float
    per: float

    def __call__(self, t: float) -> float:
        """eval_pulse, src/analysis/transient_source.rs:39-71."""
        if t < self.td:
            return self.v1
        per = self.per
        t_rel = math.fmod(t - self.td, per) if 0.0 < per < 1.7976931348623157e308 else t - self.td
        if t_rel < self.tr:
            return self.v1 + (self.v2 - self.v1) * t_rel / self.tr if self.tr > 0 else self.v2
        if t_rel < self.tr + self.pw:
            return self.v2
        if t_rel < self.tr + self.pw + self.tf:
            if self.tf > 0:
                return self.v2 + (self.v1 - self.v2) * (t_rel - self.tr - self.pw) / self.tf
            return self.v1
        return self.v1


@dataclass
class Sin:
    """SIN(vo va freq [td [theta]]) — src/ir.rs:117-123."""
    vo: float
    va: float
    freq: float
    td: float = 0.0
    theta: float = 0.0


@dataclass
class Exp:
    """EXP(v1 v2 [td1 [tau1 [td2 [tau2]]]]) — src/ir.rs:127-134; absent times
    are f64::MAX (never reached)."""
    v1: float
    v2: float
    td1: float = 0.0
    tau1: float = 1.7976931348623157e308
    td2: float = 1.7976931348623157e308
    tau2: float = 1.7976931348623157e308


@dataclass
class Pwl:
    """PWL(t1 v1 t2 v2 ...) — src/ir.rs:124-126."""
    pairs: list


def is_ground(name: str) -> bool:
    """src/compiler.rs:178-180."""
    return name == "0" or name.lower() == "gnd"


class Circuit:
    """Netlist in the reference's IR order (src/ir.rs:171-178)."""

    def __init__(self) -> None:
        self._node_ids: dict[str, int] = {}
        self.node_names: list[str] = []
        self._chunks: list[np.ndarray] = []
        self._names: list[tuple[str, int]] = []  # (prefix, count) runs or explicit names
        self._explicit_names: dict[int, str] = {}
        self._model_refs: dict[int, str] = {}
        self.n_components = 0
        self.diode_models: list[tuple[str, float, float]] = []
        self.bjt_models: list[tuple[str, float, float, float, float, float, bool]] = []
        self.mosfet_models: list[tuple[str, float, float, float, bool]] = []
        self.tran: dict[int, Pulse] = {}
        self._cached: np.ndarray | None = None

    # ---- nodes and models -------------------------------------------------
    def node(self, name: str) -> int:
        nid = self._node_ids.get(name)
        if nid is None:
            nid = len(self.node_names)
            self._node_ids[name] = nid
            self.node_names.append(name)
        return nid

    def nodes(self, names) -> np.ndarray:
        return np.fromiter((self.node(n) for n in names), dtype=np.int32)

    def fresh_nodes(self, names: list) -> np.ndarray:
        """Register names that are known to be new, in order (bulk form of node())."""
        start = len(self.node_names)
        self.node_names.extend(names)
        self._node_ids.update(zip(names, range(start, start + len(names))))
        if len(self._node_ids) != len(self.node_names):
            raise ValueError("fresh_nodes: a name was already registered")
        return np.arange(start, start + len(names), dtype=np.int32)

    def add_diode_model(self, name: str, is_: float = 1e-14, n: float = 1.0) -> None:
        self.diode_models.append((name, is_, n))

    def add_bjt_model(self, name, is_=1e-16, bf=100.0, br=1.0, nf=1.0, nr=1.0, npn=True):
        self.bjt_models.append((name, is_, bf, br, nf, nr, npn))

    def add_mosfet_model(self, name, vto=1.0, kp=2e-5, lam=0.0, nmos=True):
        self.mosfet_models.append((name, vto, kp, lam, nmos))

    def _model_index(self, table, name: str) -> int:
        for k, m in enumerate(table):  # first match by name (src/compiler.rs:416-424)
            if m[0] == name:
                return k
        return -1

    # ---- components -------------------------------------------------------
    def _append(self, kind, n0, n1, n2, values, model=-1, flags=0, ac_mag=0.0, ac_phase=0.0,
                names=None, prefix="X") -> int:
        n0 = np.atleast_1d(np.asarray(n0, dtype=np.int32))
        cnt = n0.shape[0]
        a = np.zeros(cnt, dtype=COMPONENT_DTYPE)
        a["kind"] = kind
        a["node"][:, 0] = n0
        a["node"][:, 1] = n1
        a["node"][:, 2] = n2
        a["model"] = model
        a["flags"] = flags
        a["value"] = values
        a["ac_mag"] = ac_mag
        a["ac_phase"] = ac_phase
        first = self.n_components
        self._chunks.append(a)
        if names is not None:
            for k, nm in enumerate(names):
                self._explicit_names[first + k] = nm
        else:
            self._names.append((prefix, first))
        self.n_components += cnt
        self._cached = None
        return first

    def resistor(self, name, a, b, value):
        return self._append(R, self.node(a), self.node(b), 0, value, names=[name])

    def capacitor(self, name, a, b, value):
        return self._append(C, self.node(a), self.node(b), 0, value, names=[name])

    def inductor(self, name, a, b, value):
        return self._append(L, self.node(a), self.node(b), 0, value, names=[name])

    def vsource(self, name, a, b, dc=None, ac=None, tran=None):
        flags = (HAS_DC if dc is not None else 0) | (HAS_AC if ac is not None else 0)
        mag, ph = ac if ac is not None else (0.0, 0.0)
        k = self._append(V, self.node(a), self.node(b), 0, dc or 0.0, flags=flags, ac_mag=mag,
                         ac_phase=ph, names=[name])
        if tran is not None:
            self.tran[k] = tran
        return k

    def isource(self, name, a, b, dc=None, ac=None, tran=None):
        flags = (HAS_DC if dc is not None else 0) | (HAS_AC if ac is not None else 0)
        mag, ph = ac if ac is not None else (0.0, 0.0)
        k = self._append(I, self.node(a), self.node(b), 0, dc or 0.0, flags=flags, ac_mag=mag,
                         ac_phase=ph, names=[name])
        if tran is not None:
            self.tran[k] = tran
        return k

    def diode(self, name, a, c, model):
        k = self._append(D, self.node(a), self.node(c), 0, 0.0,
                         model=self._model_index(self.diode_models, model), names=[name])
        self._model_refs[k] = model
        return k

    def bjt(self, name, c, b, e, model):
        k = self._append(Q, self.node(c), self.node(b), self.node(e), 0.0,
                         model=self._model_index(self.bjt_models, model), names=[name])
        self._model_refs[k] = model
        return k

    def mosfet(self, name, d, g, s, model):
        k = self._append(M, self.node(d), self.node(g), self.node(s), 0.0,
                         model=self._model_index(self.mosfet_models, model), names=[name])
        self._model_refs[k] = model
        return k

    # bulk forms (node ids from .node()/.nodes())
    def resistors(self, a_ids, b_ids, values, prefix="R"):
        return self._append(R, a_ids, b_ids, 0, values, prefix=prefix)

    def capacitors(self, a_ids, b_ids, values, prefix="C"):
        return self._append(C, a_ids, b_ids, 0, values, prefix=prefix)

    def inductors(self, a_ids, b_ids, values, prefix="L"):
        return self._append(L, a_ids, b_ids, 0, values, prefix=prefix)

    def diodes(self, a_ids, c_ids, model, prefix="D"):
        first = self._append(D, a_ids, c_ids, 0, 0.0,
                             model=self._model_index(self.diode_models, model), prefix=prefix)
        return first

    def isources_dc(self, a_ids, b_ids, values, prefix="I"):
        return self._append(I, a_ids, b_ids, 0, values, flags=HAS_DC, prefix=prefix)

    # ---- export -----------------------------------------------------------
    @property
    def components(self) -> np.ndarray:
        if self._cached is None:
            self._cached = (np.concatenate(self._chunks) if self._chunks
                            else np.zeros(0, dtype=COMPONENT_DTYPE))
            self._chunks = [self._cached]
        return self._cached

    def component_name(self, k: int) -> str:
        if k in self._explicit_names:
            return self._explicit_names[k]
        prefix = "X"
        for p, first in self._names:
            if first <= k:
                prefix = p
        return f"{prefix}{k}"

    def component_names(self) -> list[str]:
        return [self.component_name(k) for k in range(self.n_components)]

    def is_ground_table(self) -> np.ndarray:
        return np.array([is_ground(n) for n in self.node_names], dtype=bool)


class CCircuit:
    """ctypes view of a Circuit as the C ``ohm_circuit`` struct (kept alive here)."""

    class _S(ctypes.Structure):
        _fields_ = [
            ("n_components", ctypes.c_int32),
            ("components", ctypes.c_void_p),
            ("component_names", ctypes.POINTER(ctypes.c_char_p)),
            ("model_names", ctypes.POINTER(ctypes.c_char_p)),
            ("n_node_names", ctypes.c_int32),
            ("node_names", ctypes.POINTER(ctypes.c_char_p)),
            ("n_diode_models", ctypes.c_int32),
            ("n_bjt_models", ctypes.c_int32),
            ("n_mosfet_models", ctypes.c_int32),
            ("diode_models", ctypes.c_void_p),
            ("bjt_models", ctypes.c_void_p),
            ("mosfet_models", ctypes.c_void_p),
        ]

    def __init__(self, ckt: Circuit, with_names: bool = True) -> None:
        self.comps = np.ascontiguousarray(ckt.components)
        n = len(self.comps)
        self.node_names = (ctypes.c_char_p * max(len(ckt.node_names), 1))(
            *[s.encode() for s in ckt.node_names])
        # Names are only read by error messages; skip them for large circuits.
        if with_names and n <= 50_000:
            self.comp_names = (ctypes.c_char_p * max(n, 1))(
                *[s.encode() for s in ckt.component_names()])
        else:
            self.comp_names = None
        refs = [None] * n
        for k, m in ckt._model_refs.items():
            refs[k] = m.encode()
        self.model_names = (ctypes.c_char_p * max(n, 1))(*refs) if ckt._model_refs else None
        self.dm = np.array([(m[1], m[2]) for m in ckt.diode_models], dtype=DIODE_MODEL_DTYPE)
        self.bm = np.array([(m[1], m[2], m[3], m[4], m[5], int(m[6]), 0) for m in ckt.bjt_models],
                           dtype=BJT_MODEL_DTYPE)
        self.mm = np.array([(m[1], m[2], m[3], int(m[4]), 0) for m in ckt.mosfet_models],
                           dtype=MOSFET_MODEL_DTYPE)
        s = self._S()
        s.n_components = n
        s.components = self.comps.ctypes.data if n else None
        s.component_names = self.comp_names
        s.model_names = self.model_names
        s.n_node_names = len(ckt.node_names)
        s.node_names = self.node_names
        s.n_diode_models, s.n_bjt_models, s.n_mosfet_models = len(self.dm), len(self.bm), len(self.mm)
        s.diode_models = self.dm.ctypes.data if len(self.dm) else None
        s.bjt_models = self.bm.ctypes.data if len(self.bm) else None
        s.mosfet_models = self.mm.ctypes.data if len(self.mm) else None
        self.struct = s

    def ref(self):
        return ctypes.byref(self.struct)


class MnaC(ctypes.Structure):
    """ctypes mirror of ``ohm_mna`` (include/ohm_circuit.h)."""

    _fields_ = [
        ("size", ctypes.c_int64), ("n_nodes", ctypes.c_int64), ("n_branches", ctypes.c_int64),
        ("g_nnz", ctypes.c_int64), ("c_nnz", ctypes.c_int64),
        ("g_rowptr", ctypes.POINTER(ctypes.c_uint32)), ("g_col", ctypes.POINTER(ctypes.c_uint32)),
        ("g_val", ctypes.POINTER(ctypes.c_double)),
        ("c_rowptr", ctypes.POINTER(ctypes.c_uint32)), ("c_col", ctypes.POINTER(ctypes.c_uint32)),
        ("c_val", ctypes.POINTER(ctypes.c_double)),
        ("b_dc", ctypes.POINTER(ctypes.c_double)), ("b_ac", ctypes.POINTER(ctypes.c_double)),
        ("node_order", ctypes.POINTER(ctypes.c_int32)),
        ("branch_component", ctypes.POINTER(ctypes.c_int32)),
        ("n_diodes", ctypes.c_int64), ("n_bjts", ctypes.c_int64), ("n_mosfets", ctypes.c_int64),
        ("diodes", ctypes.c_void_p), ("bjts", ctypes.c_void_p), ("mosfets", ctypes.c_void_p),
    ]


class MnaView:
    """An ``ohm_mna`` pointing into an MnaSystem's arrays (kept alive here):
    what the entry points that take a compiled system read (G, C, b_dc and
    the device descriptors)."""

    def __init__(self, mna: "MnaSystem"):
        g, c = mna.g, mna.c
        arrs = [np.ascontiguousarray(g.row_pointers, np.uint32),
                np.ascontiguousarray(g.col_indices, np.uint32),
                np.ascontiguousarray(g.values, np.float64),
                np.ascontiguousarray(c.row_pointers, np.uint32),
                np.ascontiguousarray(c.col_indices, np.uint32),
                np.ascontiguousarray(c.values, np.float64),
                np.ascontiguousarray(mna.b_dc, np.float64)]
        descs = []
        for arr, dt in ((mna.diode_descriptors, DIODE_DESC_DTYPE),
                        (mna.bjt_descriptors, BJT_DESC_DTYPE),
                        (mna.mosfet_descriptors, MOSFET_DESC_DTYPE)):
            descs.append(np.ascontiguousarray(arr if arr is not None else np.zeros(0, dt),
                                              dtype=dt))
        self.keep = arrs + descs
        u32p, f64p = ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_double)
        m = MnaC()
        m.size, m.n_nodes = mna.size, mna.n_nodes
        m.n_branches = mna.size - mna.n_nodes
        m.g_nnz, m.c_nnz = arrs[2].shape[0], arrs[5].shape[0]
        m.g_rowptr, m.g_col = arrs[0].ctypes.data_as(u32p), arrs[1].ctypes.data_as(u32p)
        m.g_val = arrs[2].ctypes.data_as(f64p)
        m.c_rowptr, m.c_col = arrs[3].ctypes.data_as(u32p), arrs[4].ctypes.data_as(u32p)
        m.c_val = arrs[5].ctypes.data_as(f64p)
        m.b_dc = arrs[6].ctypes.data_as(f64p)
        m.n_diodes, m.n_bjts, m.n_mosfets = (len(d) for d in descs)
        m.diodes, m.bjts, m.mosfets = (d.ctypes.data if len(d) else None for d in descs)
        self.struct = m

    def ref(self):
        return ctypes.byref(self.struct)


@dataclass
class CsrMatrix:
    """CsrMatrix<T> (src/sparse.rs:9-21) with u32 indices."""
    nrows: int
    ncols: int
    values: np.ndarray
    col_indices: np.ndarray
    row_pointers: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.values.shape[0])

    def value_index(self, row: int, col: int):
        """src/sparse.rs:24-30."""
        s, e = int(self.row_pointers[row]), int(self.row_pointers[row + 1])
        hits = np.nonzero(self.col_indices[s:e] == col)[0]
        return int(s + hits[0]) if hits.size else None

    def to_dense(self) -> np.ndarray:
        d = np.zeros((self.nrows, self.ncols), dtype=self.values.dtype)
        rows = np.repeat(np.arange(self.nrows), np.diff(self.row_pointers.astype(np.int64)))
        d[rows, self.col_indices.astype(np.int64)] = self.values
        return d

    def spmv(self, x: np.ndarray) -> np.ndarray:
        rows = np.repeat(np.arange(self.nrows), np.diff(self.row_pointers.astype(np.int64)))
        y = np.zeros(self.nrows, dtype=np.result_type(self.values, x))
        np.add.at(y, rows, self.values * x[self.col_indices.astype(np.int64)])
        return y


@dataclass
class MnaSystem:
    """MnaSystem (src/compiler.rs:153-175)."""
    g: CsrMatrix
    c: CsrMatrix
    b_dc: np.ndarray
    b_ac: np.ndarray
    size: int
    node_names: list
    branch_names: list
    diode_descriptors: np.ndarray
    bjt_descriptors: np.ndarray
    mosfet_descriptors: np.ndarray
    n_nodes: int = 0
    extra: dict = field(default_factory=dict)


def mna_from_c(m: MnaC, ckt: Circuit) -> MnaSystem:
    """Copy a filled ``ohm_mna`` into numpy-owned arrays."""
    size, gn, cn = m.size, m.g_nnz, m.c_nnz

    def arr(ptr, count, dtype):
        if count == 0:
            return np.zeros(0, dtype=dtype)
        return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)

    def desc(ptr, count, dtype):
        if count == 0:
            return np.zeros(0, dtype=dtype)
        buf = (ctypes.c_char * (count * dtype.itemsize)).from_address(ptr)
        return np.frombuffer(bytes(buf), dtype=dtype).copy()

    g = CsrMatrix(size, size, arr(m.g_val, gn, np.float64), arr(m.g_col, gn, np.uint32),
                  arr(m.g_rowptr, size + 1, np.uint32))
    c = CsrMatrix(size, size, arr(m.c_val, cn, np.float64), arr(m.c_col, cn, np.uint32),
                  arr(m.c_rowptr, size + 1, np.uint32))
    b_ac = arr(m.b_ac, 2 * size, np.float64).view(np.complex128)
    order = arr(m.node_order, m.n_nodes, np.int32)
    branches = arr(m.branch_component, m.n_branches, np.int32)
    return MnaSystem(
        g=g, c=c, b_dc=arr(m.b_dc, size, np.float64), b_ac=b_ac, size=int(size),
        node_names=[ckt.node_names[i] for i in order],
        branch_names=[ckt.component_name(int(k)) for k in branches],
        diode_descriptors=desc(m.diodes, m.n_diodes, DIODE_DESC_DTYPE),
        bjt_descriptors=desc(m.bjts, m.n_bjts, BJT_DESC_DTYPE),
        mosfet_descriptors=desc(m.mosfets, m.n_mosfets, MOSFET_DESC_DTYPE),
        n_nodes=int(m.n_nodes),
        extra={"branch_components": branches},
    )


# ---------------------------------------------------------------------------
# SplitMix64 (Steele, Lea & Flood 2014) — vectorised: the k-th output of a
# stream seeded s is mix(s + (k+1)·0x9E3779B97F4A7C15 mod 2^64).
# ---------------------------------------------------------------------------
_GAMMA = np.uint64(0x9E3779B97F4A7C15)


class SplitMix64:
    def __init__(self, seed: int) -> None:
        self.seed = np.uint64(seed)
        self.pos = 0

    def next_u64(self, count: int) -> np.ndarray:
        k = np.arange(self.pos + 1, self.pos + 1 + count, dtype=np.uint64)
        self.pos += count
        with np.errstate(over="ignore"):
            z = self.seed + k * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, count: int) -> np.ndarray:
        """u = (next >> 11) · 2^-53 in [0, 1)."""
        return (self.next_u64(count) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def log_uniform(u: np.ndarray, lo: float, hi: float) -> np.ndarray:
    return lo * (hi / lo) ** u


# ---------------------------------------------------------------------------
# Benchmark circuits
# ---------------------------------------------------------------------------
def resistor_ladder(sections: int = 10_000, seed: int = 1) -> Circuit:
    """Linear DC ladder: V1 = 1 V at node 1, series 1 kΩ, log-uniform 1k–100k shunts."""
    rng = SplitMix64(seed)
    ckt = Circuit()
    ckt.vsource("V1", "1", "0", dc=1.0)
    ids = ckt.nodes(str(i) for i in range(1, sections + 2))
    ckt.resistors(ids[:-1], ids[1:], np.full(sections, 1e3), prefix="Rs")
    gnd = ckt.node("0")
    ckt.resistors(ids, np.full(sections + 1, gnd, np.int32),
                  log_uniform(rng.uniform(sections + 1), 1e3, 1e5), prefix="Rp")
    return ckt


def _mesh_resistors(ckt: Circuit, rows: int, cols: int, lo: float, hi: float,
                    rng: SplitMix64) -> np.ndarray:
    """Row-major loop: horizontal (r,c)-(r,c+1) then vertical (r,c)-(r+1,c)."""
    ids = ckt.nodes(f"n{r}_{c}" for r in range(rows) for c in range(cols)).reshape(rows, cols)
    flat = ids.ravel()
    k = np.arange(rows * cols)
    has_h = (k % cols) < cols - 1
    has_v = (k // cols) < rows - 1
    right = np.full(rows * cols, -1, dtype=np.int32)
    right[has_h] = flat[k[has_h] + 1]
    below = np.full(rows * cols, -1, dtype=np.int32)
    below[has_v] = flat[k[has_v] + cols]
    slots = np.stack([has_h, has_v], axis=1).ravel()  # per node: [h?, v?]
    a = np.repeat(flat, 2)[slots]
    b = np.stack([right, below], axis=1).ravel()[slots]
    ckt.resistors(a, b, log_uniform(rng.uniform(a.shape[0]), lo, hi))
    return ids


def diode_mesh(rows: int = 256, cols: int = 256, seed: int = 2, diode_every: int = 8) -> Circuit:
    """Nonlinear DC mesh: log-uniform 100 Ω–10 kΩ, 1 V sources at the four
    corners, a diode (IS=1e-14, N=1) from every 8th node (row-major) to ground."""
    rng = SplitMix64(seed)
    ckt = Circuit()
    ids = _mesh_resistors(ckt, rows, cols, 100.0, 1e4, rng)
    for k, (r, c) in enumerate([(0, 0), (0, cols - 1), (rows - 1, 0), (rows - 1, cols - 1)]):
        ckt.vsource(f"V{k + 1}", ckt.node_names[ids[r, c]], "0", dc=1.0)
    ckt.add_diode_model("D", 1e-14, 1.0)
    flat = ids.ravel()
    sel = flat[np.arange(0, flat.shape[0], diode_every)]
    ckt.diodes(sel, np.full(sel.shape[0], ckt.node("0"), np.int32), "D")
    return ckt


def rc_grid(rows: int = 512, cols: int = 512, n_sinks: int = 1024, seed: int = 3) -> Circuit:
    """Transient RC power grid: R log-uniform 0.1–10 Ω, node C uniform 1–100 fF,
    four 1 V VDD pads, PULSE(0 1m td 10p 10p 200p 500p) current sinks."""
    rng = SplitMix64(seed)
    ckt = Circuit()
    ids = _mesh_resistors(ckt, rows, cols, 0.1, 10.0, rng)
    flat = ids.ravel()
    gnd = ckt.node("0")
    caps = (1.0 + 99.0 * rng.uniform(flat.shape[0])) * 1e-15
    ckt.capacitors(flat, np.full(flat.shape[0], gnd, np.int32), caps)
    for k, (r, c) in enumerate([(0, 0), (0, cols - 1), (rows - 1, 0), (rows - 1, cols - 1)]):
        ckt.vsource(f"VDD{k + 1}", ckt.node_names[ids[r, c]], "0", dc=1.0)
    u = rng.uniform(2 * n_sinks).reshape(n_sinks, 2)
    nodes = np.minimum((u[:, 0] * flat.shape[0]).astype(np.int64), flat.shape[0] - 1)
    tds = u[:, 1] * 1e-9
    for k in range(n_sinks):
        ckt.isource(f"I{k + 1}", ckt.node_names[flat[nodes[k]]], "0", dc=0.0,
                    tran=Pulse(0.0, 1e-3, float(tds[k]), 10e-12, 10e-12, 200e-12, 500e-12))
    return ckt


def rlc_ladder(sections: int, rng: SplitMix64, tag: str = "") -> Circuit:
    """One AC ladder: V in 0 AC 1; per section R a_s b_s, L b_s a_{s+1}, C a_{s+1} 0."""
    ckt = Circuit()
    ckt.vsource("V1", f"a{tag}1", "0", ac=(1.0, 0.0))
    a1 = ckt.node(f"a{tag}1")
    u = rng.uniform(3 * sections).reshape(sections, 3)
    a = np.concatenate([[a1], ckt.fresh_nodes([f"a{tag}{s}" for s in range(2, sections + 2)])])
    b = ckt.fresh_nodes([f"b{tag}{s}" for s in range(1, sections + 1)])
    gnd = ckt.node("0")
    comps = np.zeros(3 * sections, dtype=COMPONENT_DTYPE)
    comps["kind"][0::3], comps["kind"][1::3], comps["kind"][2::3] = R, L, C
    comps["node"][0::3, 0], comps["node"][0::3, 1] = a[:-1], b
    comps["node"][1::3, 0], comps["node"][1::3, 1] = b, a[1:]
    comps["node"][2::3, 0], comps["node"][2::3, 1] = a[1:], gnd
    comps["model"] = -1
    comps["value"][0::3] = log_uniform(u[:, 0], 10.0, 1e4)
    comps["value"][1::3] = log_uniform(u[:, 1], 1e-9, 1e-6)
    comps["value"][2::3] = log_uniform(u[:, 2], 1e-12, 1e-9)
    ckt._chunks.append(comps)
    ckt._names.append(("X", ckt.n_components))
    ckt.n_components += comps.shape[0]
    ckt._cached = None
    return ckt


def rlc_sweep(n_circuits: int = 256, lo: int = 1000, hi: int = 64000, seed: int = 4,
              max_sections: int | None = None):
    """256 uneven RLC ladders (sections uniform lo..hi) and the 101-point
    log sweep 1 Hz–1 GHz (f_i = 10^(9i/100))."""
    rng = SplitMix64(seed)
    out = []
    for j in range(n_circuits):
        s = lo + int(rng.uniform(1)[0] * (hi - lo + 1))
        if max_sections is not None:
            s = min(s, max_sections)
        out.append(rlc_ladder(s, rng))
    freqs = 10.0 ** (9.0 * np.arange(101) / 100.0)
    return out, freqs


def amplifier_array(stages: int = 4096, seed: int = 5) -> Circuit:
    """Adjoint workload: VDD = 1.8 V; even stages common-source level-1 NMOS
    (VTO .7, KP 1e-4·W/L, LAMBDA .02), odd stages common-emitter NPN
    (IS 1e-15, BF 100), resistive gate/base dividers log-uniform 10–100 kΩ,
    drain/collector R uniform 1–10 kΩ. Each MOSFET has its own model card so
    KP·W/L is a per-device parameter."""
    rng = SplitMix64(seed)
    ckt = Circuit()
    ckt.vsource("VDD", "vdd", "0", dc=1.8)
    ckt.add_bjt_model("QN", is_=1e-15, bf=100.0)
    for j in range(stages):
        nmos = (j % 2 == 0)
        u = rng.uniform(4 if nmos else 3)
        g, d = (f"g{j}", f"d{j}")
        ckt.resistor(f"Rt{j}", "vdd", g, float(log_uniform(u[0], 1e4, 1e5)))
        ckt.resistor(f"Rb{j}", g, "0", float(log_uniform(u[1], 1e4, 1e5)))
        ckt.resistor(f"Rd{j}", "vdd", d, float(1e3 + 9e3 * u[2]))
        if nmos:
            wl = 1.0 + 9.0 * float(u[3])
            ckt.add_mosfet_model(f"NM{j}", vto=0.7, kp=1e-4 * wl, lam=0.02, nmos=True)
            ckt.mosfet(f"M{j}", d, g, "0", f"NM{j}")
        else:
            ckt.bjt(f"Q{j}", d, g, "0", "QN")
    return ckt
