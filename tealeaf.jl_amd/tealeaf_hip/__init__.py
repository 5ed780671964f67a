"""Python host interface of libtealeaf_hip, shaped like TeaLeaf.jl's.

Mirrors the reference's solver plug-in interface (src/TeaLeaf.jl:62-83,
src/settings.jl:39-78, src/chunk.jl:19-89):

    set = Settings.from_file("tea.in")       # Settings(infile)
    chunk = initialiseapp(set)               # initialiseapp!(settings)
    diffuse(chunk, set)                      # diffuse!(chunk, settings)
    set.solver.solve(chunk, set, rx, ry)     # set.solver.solve!(chunk, set, rx, ry)

Every call goes through the C ABI to the HIP implementation; nothing here
computes a field.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Callable, List, Optional

import numpy as np

from . import _lib
from ._lib import (FID, FIELDS, TL_CG, TL_CHEBY, TL_FP32_IR, TL_FP64, TL_JACOBI, TL_LAUNCH_GRAPH,
                   TL_LAUNCH_STREAM, TL_PPCG, TL_STOP_REFERENCE,
                   TL_STOP_RELATIVE, TL_WARM_GUESS, TL_WARM_LEARNED, TL_WARM_NONE, X_DENSITY, X_ENERGY,
                   X_ENERGY0, X_P, X_SD, X_U, TeaLeafError, check, load)

__all__ = [
    "Settings", "Chunk", "CG", "Cheby", "PPCG", "Jacobi", "initialiseapp", "diffuse", "fieldsummary",
    "solvefinished", "checkingvalue", "qa_check", "debugrecord", "TeaLeafError", "load", "FIELDS",
]


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class _SolverModule:
    """A solver plug-in: `solve(chunk, settings, rx, ry)` as `solve!` (TeaLeaf.jl:74)."""

    def __init__(self, name: str, code: int):
        self.name, self.code = name, code

    def solve(self, chunk: "Chunk", settings: "Settings", rx: float, ry: float) -> dict:
        p = settings.params(solver=self.code)
        st = _lib.SolveStats()
        check(load().tl_solve(chunk._ctx, C.byref(p), rx, ry, C.byref(st)), chunk._ctx, "solve")
        out = st.as_dict()
        chunk._record_solve(p, out)
        return out

    def __repr__(self):
        return f"<solver {self.name}>"


CG = _SolverModule("cg", TL_CG)
Cheby = _SolverModule("cheby", TL_CHEBY)
PPCG = _SolverModule("ppcg", TL_PPCG)
Jacobi = _SolverModule("jacobi", TL_JACOBI)
_BY_CODE = {m.code: m for m in (CG, Cheby, PPCG, Jacobi)}
SOLVERS = {m.name: m for m in (CG, Cheby, PPCG, Jacobi)}


class Settings:
    """Parsed tea.in deck (settings.jl:39-78), parsed by the library's own parser."""

    def __init__(self, deck: _lib.Deck, text: str = ""):
        self._deck = deck
        self.text = text
        self.solver = _BY_CODE[deck.params.solver]
        self.precision = deck.params.precision
        self.stop_rule = deck.params.stop_rule
        self.rel_tol = deck.params.rel_tol
        self.warm_start = deck.params.warm_start
        self.ir_delta = deck.params.ir_delta
        self.debugfile = ""  # Settings.debugfile (settings.jl:77, run.jl --debug-out)

    @classmethod
    def from_text(cls, text: str) -> "Settings":
        d = _lib.Deck()
        rc = load().tl_deck_parse(text.encode(), C.byref(d))
        if rc != _lib.TL_OK:
            raise TeaLeafError(rc, "could not parse deck")
        return cls(d, text)

    @classmethod
    def from_file(cls, path: str) -> "Settings":
        with open(path) as f:
            return cls.from_text(f.read())

    # reference field names
    xcells = property(lambda s: s._deck.x_cells)
    ycells = property(lambda s: s._deck.y_cells)
    halodepth = property(lambda s: s._deck.halo_depth)
    xmin = property(lambda s: s._deck.xmin)
    ymin = property(lambda s: s._deck.ymin)
    xmax = property(lambda s: s._deck.xmax)
    ymax = property(lambda s: s._deck.ymax)
    dx = property(lambda s: s._deck.dx)
    dy = property(lambda s: s._deck.dy)
    endstep = property(lambda s: s._deck.end_step)
    dtinit = property(lambda s: s._deck.dt_init)
    summaryfrequency = property(lambda s: s._deck.summary_frequency)
    checkresult = property(lambda s: bool(s._deck.check_result))
    coefficient = property(lambda s: s._deck.params.coefficient)
    maxiters = property(lambda s: s._deck.params.max_iters)
    eps = property(lambda s: s._deck.params.eps)
    presteps = property(lambda s: s._deck.params.presteps)
    ppcginnersteps = property(lambda s: s._deck.params.ppcg_inner_steps)
    n_warnings = property(lambda s: s._deck.n_warnings)

    @property
    def states(self) -> List[dict]:
        out = []
        for i in range(self._deck.nstates):
            s = self._deck.states[i]
            out.append({n: getattr(s, n) for n, _ in s._fields_})
        return out

    def params(self, **over) -> _lib.SolveParams:
        p = _lib.SolveParams()
        C.memmove(C.byref(p), C.byref(self._deck.params), C.sizeof(p))
        p.solver = self.solver.code
        p.precision = self.precision
        p.stop_rule = self.stop_rule
        p.rel_tol = self.rel_tol
        p.warm_start = self.warm_start
        p.ir_delta = self.ir_delta
        for k, v in over.items():
            setattr(p, k, v)
        return p

    def set(self, **kw):
        """Override deck values (max_iters=..., presteps=..., eps=...)."""
        for k, v in kw.items():
            if k == "solver":
                self.solver = SOLVERS[v] if isinstance(v, str) else v
            elif k in ("precision", "stop_rule", "rel_tol", "warm_start", "ir_delta", "debugfile"):
                setattr(self, k, v)
            else:
                setattr(self._deck.params, k, v)
        return self


class Chunk:
    """Device-resident field store (chunk.jl:19-89) owned by one tl_ctx."""

    def __init__(self, settings: Settings, device: int = 0):
        L = load()
        ctx = C.c_void_p()
        d = settings._deck
        check(L.tl_create(C.byref(ctx), d.x_cells, d.y_cells, d.halo_depth, d.xmin, d.ymin, d.dx, d.dy, device),
              None, "tl_create")
        self._ctx = ctx
        X, Y = C.c_int(), C.c_int()
        L.tl_dims(ctx, C.byref(X), C.byref(Y))
        self.X, self.Y, self.hd = X.value, Y.value, d.halo_depth
        self._init_host_state(settings)

    def _init_host_state(self, settings: Settings):
        """The reference Chunk's members that are not device fields
        (chunk.jl:40-60), built as Chunk(settings) builds them (chunk.jl:68-89)
        and kept on the host for debugrecord: mesh vectors, volume and areas,
        the Chebyshev scalars and the CG / Chebyshev coefficient histories
        (length maxiters), which solves update as the reference's solve! does."""
        d = settings._deck
        x, y, h = self.X, self.Y, self.hd
        self.vertexx = d.xmin + d.dx * (np.arange(1, x + 2) - h - 1)
        self.vertexy = d.ymin + d.dy * (np.arange(1, y + 2) - h - 1)
        self.cellx = 0.5 * (self.vertexx[:-1] + self.vertexx[1:])
        # chunk.jl:44 adds vertexx to vertexy (as written, reproduced here);
        # with x != y that broadcast fails in the reference, and vertexy is used
        self.celly = 0.5 * (self.vertexy[:-1] + (self.vertexx if x == y else self.vertexy)[1:])
        self.volume = np.full((y, x), d.dx * d.dy)
        self.xarea = np.full((y, x), d.dy)
        self.yarea = np.full((y, x), d.dx)
        self.theta = self.eigmin = self.eigmax = 0.0
        n = d.params.max_iters
        self.cgalpha, self.cgbeta = np.zeros(n), np.zeros(n)
        self.chebyalpha, self.chebybeta = np.zeros(n), np.zeros(n)

    def _record_solve(self, p: _lib.SolveParams, st: dict):
        """What an FP64 solve! leaves in the reference's chunk besides the
        fields: cgα/cgβ of its CG iterations (CG.jl:37-38), and for
        Chebyshev/PPCG after the switch eigmin/eigmax (kernels.jl:47-48), θ and
        chebyα/chebyβ (Cheby.coef!, Cheby.jl:121-135; PPCG.jl:42).  The
        mixed-precision solver has no counterpart and records nothing."""
        if p.precision != TL_FP64:
            return
        n = len(self.cgalpha)
        if p.solver in (TL_CG, TL_CHEBY, TL_PPCG):
            k = min(int(st["cg_iters"]), n)
            if k > 0:
                a, b = self.cg_history(k)
                self.cgalpha[:k], self.cgbeta[:k] = a, b
        if p.solver in (TL_CHEBY, TL_PPCG) and st["eigmax"] > 0.0:
            self.eigmin, self.eigmax = st["eigmin"], st["eigmax"]
            m = p.max_iters - int(st["cg_iters"]) if p.solver == TL_CHEBY else p.ppcg_inner_steps
            m = max(0, min(m, n))
            a, b, th = np.zeros(max(m, 1)), np.zeros(max(m, 1)), C.c_double()
            load().tl_cheby_coef(self.eigmin, self.eigmax, m, _dp(a), _dp(b), C.byref(th))
            self.chebyalpha[:m], self.chebybeta[:m] = a[:m], b[:m]
            self.theta = th.value

    def close(self):
        if getattr(self, "_ctx", None):
            load().tl_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        self.close()

    def field(self, name: str) -> np.ndarray:
        """(Y, X) array: [j, k] is the reference's [k+1, j+1]."""
        a = np.empty((self.Y, self.X))
        check(load().tl_get_field(self._ctx, FID[name], _dp(a)), self._ctx, "tl_get_field")
        return a

    def interior(self, name: str) -> np.ndarray:
        h = self.hd
        return self.field(name)[h:self.Y - h, h:self.X - h]

    def set_field(self, name: str, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == (self.Y, self.X)
        check(load().tl_set_field(self._ctx, FID[name], _dp(a)), self._ctx, "tl_set_field")

    def set_guess(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == (self.Y, self.X)
        check(load().tl_set_guess(self._ctx, _dp(a)), self._ctx, "tl_set_guess")

    def halo_update(self, mask: int, depth: int = 1):
        check(load().tl_halo_update(self._ctx, mask, depth), self._ctx, "tl_halo_update")

    def apply_operator(self, name: str, rx: float, ry: float, coefficient: int = 1) -> np.ndarray:
        check(load().tl_apply_operator(self._ctx, FID[name], rx, ry, coefficient), self._ctx, "tl_apply_operator")
        return self.field("w")

    def cg_history(self, n: int):
        a, b = np.zeros(n), np.zeros(n)
        check(load().tl_get_cg_history(self._ctx, _dp(a), _dp(b), n), self._ctx, "tl_get_cg_history")
        return a, b


def initialiseapp(settings: Settings, device: int = 0) -> Chunk:
    """initialiseapp!, TeaLeaf.jl:35-44 (rasterisation on the device)."""
    chunk = Chunk(settings, device)
    d = settings._deck
    check(load().tl_initialise(chunk._ctx, d.states, d.nstates), chunk._ctx, "tl_initialise")
    return chunk


def solvefinished(chunk: Chunk, settings: Settings):
    """solvefinished!, kernels.jl:166-170."""
    check(load().tl_solve_finished(chunk._ctx, int(settings.checkresult)), chunk._ctx, "tl_solve_finished")


def fieldsummary(chunk: Chunk) -> np.ndarray:
    """fieldsummary, kernels.jl:119-133: [vol, mass, ie, temp]."""
    out = np.zeros(4)
    check(load().tl_field_summary(chunk._ctx, _dp(out)), chunk._ctx, "tl_field_summary")
    return out


def checkingvalue(settings: Settings, problemfile: str = "tea.problems") -> float:
    """checkingvalue, settings.jl:180-196: the problem file's value for
    (xcells, ycells, endstep); 1.0 with a warning when no line matches."""
    v, found = C.c_double(), C.c_int()
    rc = load().tl_checking_value(problemfile.encode(), settings.xcells, settings.ycells, settings.endstep,
                                  C.byref(v), C.byref(found))
    if rc != _lib.TL_OK:
        raise TeaLeafError(rc, f"checkingvalue: cannot read {problemfile}")
    if not found.value:
        import warnings
        warnings.warn("Problem was not found in the test problems file.")
    return v.value


def qa_check(chunk: Chunk, settings: Settings, problemfile: str = "tea.problems") -> Optional[dict]:
    """The QA half of fieldsummary, kernels.jl:119-133: temp = sum vol*rho*u
    against the checking value; None when check_result is off."""
    if not settings.checkresult:
        return None
    actual = float(fieldsummary(chunk)[3])
    cv = checkingvalue(settings, problemfile)
    d = C.c_double()
    passed = bool(load().tl_qa_check(actual, cv, C.byref(d)))
    return {"actual": actual, "checking_value": cv, "qa_diff": d.value, "passed": passed}


def _julia_float(x: float) -> str:
    """A Float64 as Julia's print writes it: the shortest round-trip digits,
    with "e" notation outside 1e-4 <= |x| < 1e6 ("1.0e-5", "1.234567e6"),
    ".0" on integral values and "NaN"/"Inf" spellings."""
    if x != x:
        return "NaN"
    if x in (float("inf"), float("-inf")):
        return "Inf" if x > 0 else "-Inf"
    if x == 0.0:
        return "-0.0" if str(x).startswith("-") else "0.0"
    ax = abs(x)
    if 1e-4 <= ax < 1e6:
        r = repr(x)
        if "e" in r:  # Python switches to exponents earlier for small values
            r = np.format_float_positional(x, unique=True, trim="-")
        return r if ("." in r) else r + ".0"
    m, e = np.format_float_scientific(x, unique=True, trim="-").split("e")
    if "." not in m:
        m += ".0"
    return f"{m}e{int(e)}"


# the members of the reference's Chunk in declaration order (chunk.jl:19-60),
# as debugrecord writes them (x and y excluded), and where this host keeps them
DEBUG_MEMBERS = (list(FIELDS) + ["vertexx", "vertexy", "cellx", "celly", "volume", "xarea", "yarea",
                                 "\u03b8", "eigmin", "eigmax", "cg\u03b1", "cg\u03b2", "cheby\u03b1",
                                 "cheby\u03b2"])
_HOST_MEMBER = {"\u03b8": "theta", "cg\u03b1": "cgalpha", "cg\u03b2": "cgbeta", "cheby\u03b1": "chebyalpha",
                "cheby\u03b2": "chebybeta"}


def _julia_string(v) -> str:
    """getstring, TeaLeaf.jl:104-107: a matrix as one line per column (the
    values of k = 1..X separated by spaces), a vector as one line, a scalar
    as Julia's string()."""
    if np.ndim(v) == 0:
        return _julia_float(float(v))
    if np.ndim(v) == 1:
        return " ".join(_julia_float(float(x)) for x in v)
    return "\n".join(" ".join(_julia_float(float(x)) for x in row) for row in v)


def debugrecord(chunk: Chunk, path: str):
    """debugrecord, TeaLeaf.jl:90-102: append every member of the chunk
    except x and y to `path`, in declaration order (chunk.jl:19-60): the
    member name, its getstring (TeaLeaf.jl:104-107), then an empty line; two
    newlines after the chunk.  Device fields are read back; the other members
    come from the host mirror the solves keep (Chunk._record_solve).  The
    number format follows Julia's print; it is not checked against Julia
    (none here)."""
    with open(path, "a", encoding="utf-8") as f:
        for name in DEBUG_MEMBERS:
            v = chunk.field(name) if name in FIELDS else getattr(chunk, _HOST_MEMBER.get(name, name))
            f.write(name + "\n")
            f.write(_julia_string(v))
            f.write("\n\n")
        f.write("\n\n")


def step(chunk: Chunk, settings: Settings) -> dict:
    """One timestep of diffuse!, TeaLeaf.jl:67-77, through tl_step."""
    p = settings.params()
    st = _lib.SolveStats()
    check(load().tl_step(chunk._ctx, C.byref(p), settings.dtinit, C.byref(st)), chunk._ctx, "tl_step")
    out = st.as_dict()
    chunk._record_solve(p, out)
    return out


def diffuse(chunk: Chunk, settings: Settings, nsteps: Optional[int] = None,
            callback: Optional[Callable[[int, np.ndarray, dict], None]] = None):
    """diffuse!, TeaLeaf.jl:62-83: returns (summaries[n, 4], stats list).
    With settings.debugfile set, the file is replaced and the chunk is dumped
    to it at the start of every timestep (TeaLeaf.jl:63-68)."""
    n = int(nsteps if nsteps is not None else settings.endstep)
    if settings.debugfile and os.path.isfile(settings.debugfile):
        os.remove(settings.debugfile)
    summ, stats = [], []
    for t in range(1, n + 1):
        if settings.debugfile:
            debugrecord(chunk, settings.debugfile)
        st = step(chunk, settings)
        s = fieldsummary(chunk)
        summ.append(s)
        stats.append(st)
        if callback:
            callback(t, s, st)
    return np.array(summ), stats
