"""Reader for the reference's input decks (the subset the mef Biot u-p path
runs) and the macro program they describe.

The reference reads a deck with ``readmacro`` (``Rdata.f:1945-2001``): a
line is split into blank-separated words, the first word of a line selects a
macro, and data lines are read word by word.  The top level
(``Main.f:405-470``) runs macros in order, ``loop n ... next`` repeating the
macros between them; the ``mesh`` macro reads the mesh block with
``rdat_pm`` (``Rdata.f:1-1037``).  This module reads the same words into a
:class:`~poromec_amd.problem.Deck` plus the initial state, the solver
settings and the list of time-step and print macros; ``poromec_amd.mef``
runs that program on the GPU and writes the reference's output files.

Macros read (top level): ``config`` (``reord``, ``bvtk``, ``legacy_vtk``),
``solver``, ``gravity``, ``nl``, ``conseq``, ``dt``, ``mesh``, ``setpnode``,
``setprint``, ``pnup``, ``pnsf``, ``pres``, ``solv``, ``loop``/``next``,
``stop``.  Mesh block: the parameter line, ``materials``, ``fmaterials``,
``hexa8``, ``coordinates``, ``constrainpmec``, ``elmtloads``,
``elmtpresloads``, ``loads`` (types 1-8 inline, 39-43 from a file),
``initialpres``, ``hydrostatic``, ``insert``/``return``, ``end``.

Earlier forms found in the decks the reference ships with its outputs are
also read, each where the current reader would reject or misread it:

* ``config`` and ``gravity`` written inline (``config`` ... ``end config``;
  ``gravity gx gy gz``, with the modulus taken as the vector's norm as
  ``Main.f:320-323`` does for the default) instead of naming a file;
* ``pcg maxit tol`` and ``maxnlit n`` in place of ``solver <file>`` and
  ``nl``; a zero tolerance becomes the machine epsilon as the ``solver``
  reader does (``Rdata.f:3196``);
* ``loads`` entries of type 39-43 given inline as ``nc type n v1 .. vn``,
  a load constant in time;
* element type 7 for the hexa20/8 Biot element (17 in the current library)
  in a ``maxno 20`` mesh; a ``maxno 8`` mesh (hexa8 kept as read) is the
  equal-order Q1/Q1 element, type 7 -- the ``ELMT07_PM`` the reference
  reserves without implementing it (``Elmlib03.f:22-23``);
* blocks the current reader does not know (``faceR1`` ... ``end faceR1``):
  their ``end <name>`` line closes that block and not the mesh block.

Macros that only write post-processing files this reader does not produce
(``pgeo``, ``deltatc``, ``setpi``, ...) are recorded in
:attr:`MefInput.skipped` and otherwise ignored; macros that would change
the system this path solves (initial stresses, nodal forces, hexa20 input,
plasticity, ...) raise :class:`DeckError`.
"""
from __future__ import annotations

import dataclasses
import math
import os
from dataclasses import dataclass, field

import numpy as np

from .problem import NDF, NEN, NPAR, NTERMS, NUMLOAD, PROP, Deck, empty_load_table
from .solver import SolverConfig
from .vtk import DEFAULT_FLAGS, PRINT_WORDS

# element type ids accepted: the hexa20/8 Biot element (elmt17_pm; 7 in the
# shipped decks) and its variable-property form (elmt18_pm)
ELEMENT_TYPES = (17, 7, 18)
# post-processing macros: they write files this path does not produce
OUTPUT_MACROS = ("pgeo", "setpi", "ppi", "pcoo", "pcolors", "pmecres", "ptermres", "deltatc")
# macros that change the problem beyond this path
UNSUPPORTED_TOP = ("block_pu", "hextotet", "solvt", "solvm", "presolv", "neqs", "run")
UNSUPPORTED_MESH = ("hexa20", "bar2", "tria3", "quad4", "tetra4", "tetra10", "constraindisp",
                    "nodalforces", "nodalloads", "constraintemp", "nodalthermloads",
                    "hydrostress", "initialstress", "parallel", "tria3ov", "quad4ov",
                    "tetra4ov", "hexa8ov", "tetra10ov", "hexa20ov")
SOLVER_WORDS = {"cg": "pcg", "pcg": "pcg", "sqmr": "sqmr", "minres": "minres",
                "bicgs": "bicgstab", "bicgstab": "bicgstab", "gmres": "gmres"}
EPS = 2.0 ** -53  # smachn(): the machine epsilon the reference substitutes for tol 0


class DeckError(ValueError):
    """A deck this path cannot read (the reference stops in ``stop_mef``)."""


def fnum(word: str) -> float:
    """A Fortran list-directed real: ``1.d-3``, ``0.3369d+00``, ``-0.0d0``."""
    try:
        return float(word.replace("d", "e").replace("D", "e"))
    except ValueError:
        raise DeckError(f"not a number: {word!r}") from None


def inum(word: str) -> int:
    try:
        return int(word)
    except ValueError:
        raise DeckError(f"not an integer: {word!r}") from None


def flag(word: str) -> bool:
    w = word.strip(".").lower()
    if w in ("true", "t"):
        return True
    if w in ("false", "f"):
        return False
    raise DeckError(f"not a logical: {word!r}")


class _Lines:
    """Lines of a deck as word lists, following ``insert`` files.

    A carriage return ends a word as a blank does (``Rdata.f:1985-1987``).
    An included file ends at its ``return`` line or at its end."""

    def __init__(self, path):
        self.stack = []
        self.push(path)

    def push(self, path):
        base = os.path.dirname(self.stack[-1][0]) if self.stack else ""
        full = path if os.path.isabs(path) or not base else os.path.join(base, path)
        try:
            with open(full, "r", errors="replace") as f:
                text = f.read()
        except OSError:
            raise DeckError(f"File {path} not found !") from None
        self.stack.append([full, text.replace("\r", " ").split("\n"), 0])

    def pop(self):
        if len(self.stack) > 1:
            self.stack.pop()

    def path(self, name):
        base = os.path.dirname(self.stack[-1][0])
        return name if os.path.isabs(name) else os.path.join(base, name)

    def next(self):
        """Words of the next line; None at the end of the main file."""
        while self.stack:
            full, lines, pos = self.stack[-1]
            if pos >= len(lines):
                if len(self.stack) == 1:
                    return None
                self.stack.pop()
                continue
            self.stack[-1][2] = pos + 1
            return lines[pos].split()
        return None

    def where(self):
        full, _, pos = self.stack[-1]
        return f"{os.path.basename(full)}:{pos}"


def _file_words(path):
    with open(path, "r", errors="replace") as f:
        return f.read().replace("\r", " ").split()


@dataclass
class MefInput:
    """Everything a deck asks of the mef path."""

    deck: Deck
    u0: np.ndarray                 # (nnodev, 4) initial values of the vertex dofs
    cfg: SolverConfig
    reorder: bool
    pnodes: list                   # nodes of setpnode (1-based)
    program: list                  # ("solv",) / ("pnup",) / ("pnsf",) / ("pres",) / ("dt", value)
    skipped: list = field(default_factory=list)
    element_type: int = 17
    path: str = ""
    ie: np.ndarray = None          # (numat) element type per material: 17 or 18, or 7 (maxno 8)
    vprop: tuple = (False, False, False, False)  # conseq: vprop, kozeny_carman, density, hashin_shtrikman
    print_flags: tuple = DEFAULT_FLAGS  # setprint (print_flag 1-13) as the pres macro uses it
    vtk_legacy: bool = False       # config legacy_vtk: .vtk instead of .vtu
    vtk_binary: bool = False       # config bvtk


def read_deck(path) -> MefInput:
    """Read a deck and every file it includes."""
    src = _Lines(str(path))
    cfg = SolverConfig(solver="sqmr", precond=2, tol=1.0e-11, maxit=50000, ngram=150,
                       maxnlit=500, nltol=1.0e-4, stop_crit=2)  # Main.f:262-289 defaults
    st = dict(dt=1.0, gravity=np.array([0.0, 0.0, -9.81]), reorder=True, mesh=None,
              pnodes=[], skipped=[], vprop=[False] * 4, print_flags=list(DEFAULT_FLAGS),
              vtk_legacy=False, vtk_binary=False)
    st["gravity_mod"] = float(np.linalg.norm(st["gravity"]))
    program = []
    loop = None  # (count, [ops]) while inside loop ... next
    while True:
        w = src.next()
        if w is None:
            break
        if not w:
            continue
        mc = w[0][:8]
        ops = loop[1] if loop is not None else program
        if mc == "loop":
            if len(w) < 2:
                raise DeckError("Error reading macro (LOOP) !")
            loop = (inum(w[1]), [])
        elif mc == "next":
            if loop is None:
                continue
            program.extend(loop[1] * loop[0])
            loop = None
        elif mc in ("solv", "pnup", "pnsf", "pres"):
            ops.append((mc,))
        elif mc == "dt":
            if len(w) < 2:
                raise DeckError("Error reading macro (DT) !")
            ops.append(("dt", fnum(w[1])))
        elif mc == "stop":
            break
        elif mc == "config":
            _read_config(src, w, st)
        elif mc == "gravity":
            _read_gravity(src, w, st)
        elif mc == "conseq":
            if st["mesh"] is not None:
                raise DeckError("This macro can only be used before macro mesh")
            st["vprop"] = _read_conseq(src.path(w[1]) if len(w) > 1 else "")
        elif mc == "solver":
            if st["mesh"] is not None:
                raise DeckError("This macro can only be used before macro mesh")
            _read_solver_file(src.path(w[1]) if len(w) > 1 else "", cfg)
        elif mc in ("pcg", "cg", "sqmr", "minres", "bicgstab") and len(w) >= 3:
            cfg.solver = SOLVER_WORDS[mc]
            cfg.maxit = inum(w[1])
            cfg.tol = fnum(w[2]) or EPS
        elif mc == "nl":
            if len(w) < 4:
                raise DeckError("Error reading macro (NL) !")
            cfg.maxnlit, cfg.nltol, cfg.stop_crit = inum(w[1]), fnum(w[2]), inum(w[3])
        elif mc == "maxnlit":
            # the earlier loop: residual test |R|/|R0| and the step kept
            # after maxnlit iterations (the decks' own outputs need both)
            cfg.maxnlit, cfg.stop_crit, cfg.maxnlit_accept = inum(w[1]), 1, True
        elif mc == "setpnode":
            st["pnodes"] = _read_pnode(src.path(w[1]))
        elif mc == "setprint":
            if len(w) < 2:
                raise DeckError("setprint without a file")
            _read_setprint(src.path(w[1]), st)
        elif mc == "mesh":
            st["mesh"] = _read_mesh(src, st)
            # rdat_pm fixes the printed nodes (print_quad); an elastic run
            # prints no consolidation pressure nor plastic flag (Main.f:552-556)
            st["print_flags"][9] = st["print_flags"][10] = False
            st["quad"] = st["print_flags"][0]
        elif mc in OUTPUT_MACROS:
            st["skipped"].append(mc)
        elif mc in UNSUPPORTED_TOP:
            raise DeckError(f"macro {mc!r} is outside the mef Biot u-p path")
        # any other word is skipped, as the reference's macro loop does
    if st["mesh"] is None:
        raise DeckError("deck has no mesh macro")
    if cfg.stop_crit not in (1, 2):
        raise DeckError(f"nonlinear stop criterion {cfg.stop_crit} (only 1 and 2 are in this path)")
    m = st["mesh"]
    deck = Deck(0, 0, 0, m["ix"], m["x"], m["idv"], m["e"], m["eload"], m["eloadp"], m["load"],
                m["fload"], gravity=m["gravity"], gravity_mod=m["gravity_mod"], dt=st["dt"],
                nsteps=sum(1 for op in program if op[0] == "solv"), nen=m["nen"])
    dts = [op[1] for op in program if op[0] == "dt"]
    if dts:
        deck.dt = dts[0]
    if cfg.solver == "pcg" and cfg.precond not in (1, 2, 5):
        raise DeckError(f"preconditioner {cfg.precond} is outside this path")
    if m["nen"] == 8:  # maxno 8: every material the Q1/Q1 element, no conseq
        if (m["ie"] != 7).any() or st["vprop"][0]:
            raise DeckError("maxno 8: every material must be element type 7 (Q1/Q1), "
                            "without conseq vprop")
    ie = np.where(m["ie"] == 7, 7 if m["nen"] == 8 else 17, m["ie"]).astype(np.int32)
    any18 = bool((ie == 18).any())
    if st["vprop"][0] != any18:  # check_element, Rdata.f:3744-3758
        raise DeckError("Erro: Invalid elements!! (elmt18 materials go with conseq vprop, "
                        "elmt17 without it)")
    if st["vtk_binary"] and not st["vtk_legacy"] and any(op[0] == "pres" for op in program):
        raise DeckError("bvtk without legacy_vtk: the reference has no binary VTU writer")
    return MefInput(deck=deck, u0=m["u0"], cfg=cfg, reorder=st["reorder"], pnodes=st["pnodes"],
                    program=program, skipped=st["skipped"], element_type=m["etype"],
                    path=str(path), ie=ie, vprop=tuple(st["vprop"]),
                    print_flags=tuple(st["print_flags"]), vtk_legacy=st["vtk_legacy"],
                    vtk_binary=st["vtk_binary"])


# -- top-level macros ----------------------------------------------------------------
def _read_config(src, w, st):
    """``config`` (read_config, Rdata.f:2929-3130): ``reord`` is the one key
    this path uses; the OpenMP, memory and VTK keys are read and ignored."""
    if len(w) > 1:  # config <file>
        words = _file_words(src.path(w[1]))
    else:  # the earlier inline block
        words = []
        while True:
            line = src.next()
            if line is None:
                raise DeckError("config block without end")
            if line and line[0] == "end":
                break
            words.extend(line)
    i = 0
    while i < len(words) and words[i] != "end":
        key = words[i]
        if key in ("memory", "omp_elmt", "omp_solver", "nth_elmt", "nth_solver", "reord", "bvtk",
                   "legacy_vtk", "newton_raphson"):
            if i + 1 >= len(words):
                raise DeckError(f"config: {key} without a value")
            if key == "reord":
                st["reorder"] = flag(words[i + 1])
            elif key == "bvtk":
                st["vtk_binary"] = flag(words[i + 1])
            elif key == "legacy_vtk":
                st["vtk_legacy"] = flag(words[i + 1])
            elif key == "newton_raphson" and flag(words[i + 1]):
                raise DeckError("newton_raphson (tangent rebuilt every iteration) is outside this path")
            i += 2
        else:
            i += 1


def _read_gravity(src, w, st):
    """``gravity <file>`` with gx gy gz |g| (read_gravity, Rdata.f:3543-3581),
    or the earlier inline ``gravity gx gy gz``."""
    if len(w) >= 4:
        g = np.array([fnum(v) for v in w[1:4]])
        mg = float(math.sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]))
    elif len(w) == 2:
        words = _file_words(src.path(w[1]))
        if len(words) < 4:
            raise DeckError("Erro na leitura da macro (GRAVITY)")
        g = np.array([fnum(v) for v in words[:3]])
        mg = fnum(words[3])
    else:
        raise DeckError("Erro na leitura da macro (GRAVITY)")
    st["gravity"], st["gravity_mod"] = g, mg


def _read_setprint(path, st):
    """``setprint <file>``: the fields the pres macro writes, the first word
    of each line until ``end`` (set_print_vtk_pm, Rdata.f:3383-3520); the
    file clears every flag first."""
    try:
        with open(path, "r", errors="replace") as f:
            lines = f.read().replace("\r", " ").split("\n")
    except OSError:
        raise DeckError(f"File {path} not found !") from None
    flags = [False] * len(DEFAULT_FLAGS)
    for ln in lines:
        w = ln.split()
        if not w:
            continue
        if w[0] == "end":
            break
        if w[0] in PRINT_WORDS:
            flags[PRINT_WORDS.index(w[0])] = True
    if st["mesh"] is not None:
        if flags[0] != st["quad"]:
            raise DeckError("setprint quadratic after the mesh macro (the printed nodes are "
                            "fixed when the mesh is read)")
        if flags[9] or flags[10]:
            raise DeckError("pconsolidation / elplastic are fields of the plastic path")
    st["print_flags"] = flags


def _read_conseq(path):
    """``conseq <file>``: the constitutive options, one word a line until
    ``end`` (read_constitutive_equation, Rdata.f:2545-2640); any of
    kozeny_carman, density, hashin_shtrikman turns vprop on."""
    try:
        words = _file_words(path)
    except OSError:
        raise DeckError(f"File {path} not found !") from None
    flags = [False] * 4
    names = ("vprop", "kozeny_carman", "density", "hashin_shtrikman")
    for wd in words:
        if wd == "end":
            break
        if wd == "plastic":
            raise DeckError("plasticity (conseq plastic) is outside the mef Biot u-p path")
        if wd in names:
            flags[names.index(wd)] = True
    if any(flags[1:]):
        flags[0] = True
    return flags


def _read_solver_file(path, cfg):
    """The ``solver`` config file (read_solver_config_pm, Rdata.f:3141-3305;
    names: set_solver, Solv.f:2683; set_precond, Precond.f:1476)."""
    try:
        words = _file_words(path)
    except OSError:
        raise DeckError(f"File {path} not found !") from None
    precs = {"none": 1, "diag": 2, "ildlt": 3, "illt": 4, "diagm": 5, "bdiag": 6, "diags": 7}
    i = 0
    while i < len(words) and words[i] != "end":
        key, val = words[i], words[i + 1] if i + 1 < len(words) else None
        if val is None:
            raise DeckError(f"*** Error reading macro: {key} in macro solver!")
        if key == "name":
            if val not in SOLVER_WORDS:
                raise DeckError(f"solver {val!r} is outside this path")
            cfg.solver = SOLVER_WORDS[val]
        elif key == "it":
            cfg.maxit = inum(val)
        elif key == "tol":
            cfg.tol = fnum(val) or EPS
        elif key == "precond":
            if precs.get(val) not in (1, 2, 5):
                raise DeckError(f"preconditioner {val!r} is outside this path")
            cfg.precond = precs[val]
        elif key == "nkrylov":
            cfg.ngram = inum(val)
        elif key in ("histlog", "alfap", "alfau", "ctol", "cmaxit", "fprint"):
            pass
        else:
            i += 1
            continue
        i += 2


def _read_pnode(path):
    """``setpnode`` file: the count, then one node per line (readpnode,
    Rdata.f:3916-3955); at most 50 nodes."""
    try:
        words = _file_words(path)
    except OSError:
        raise DeckError(f"File {path} not found !") from None
    n = inum(words[0])
    if n > 50:
        raise DeckError("number max print node is 50 !!!")
    return [inum(v) for v in words[1:1 + n]]


# -- the mesh block --------------------------------------------------------------------
def _read_mesh(src, st):
    par = _parameters(src)
    nnode, numel, numat = par["nnode"], par["numel"], par["numat"]
    if par["ndf"] != NDF or par["dim"] != 3:
        raise DeckError(f"ndf {par['ndf']} dim {par['dim']}: this path is the 3-D u-p system "
                        "(ndf 4, dim 3)")
    if par["maxno"] not in (8, NEN):
        raise DeckError(f"maxno {par['maxno']}: this path is the hexa20/8 element (maxno 20) "
                        "or the Q1/Q1 element (maxno 8)")
    m = dict(ix=np.zeros((numel, NEN + 1), np.int32), x=np.zeros((nnode, 3)),
             idv=np.zeros((nnode, NDF), np.int32), e=np.zeros((numat, PROP)),
             ie=np.zeros(numat, np.int32), eload=np.zeros((numel, 7), np.int32),
             eloadp=np.zeros((numel, 7), np.int32), u0=np.zeros((nnode, NDF)),
             gravity=st["gravity"].copy(), gravity_mod=st["gravity_mod"], etype=17,
             nen=par["maxno"])
    m["load"], m["fload"] = empty_load_table()
    have_el = have_x = False
    while True:
        w = src.next()
        if w is None:
            raise DeckError("mesh block without end")
        if not w:
            continue
        mc = w[0][:15]
        if mc == "end":
            if len(w) == 1 or w[1] == "mesh":
                break
            continue  # closes a block this reader does not know
        if mc == "materials":
            _materials(src, m, numat)
        elif mc == "fmaterials":
            _fmaterials(src, m, numat)
        elif mc == "hexa8":
            _elconn(src, m["ix"], numel, 8)
            have_el = True
        elif mc == "coordinates":
            _coord(src, m["x"], nnode)
            have_x = True
        elif mc in ("constrainpmec", "elmtloads", "elmtpresloads", "initialpres"):
            if not have_el:
                raise DeckError(f"MACRO: {mc} !! Unread Elements")
            if mc == "constrainpmec":
                _bound(src, m["idv"], nnode, NDF, "restricoes")
            elif mc == "elmtloads":
                _bound(src, m["eload"], numel, 7, "cargas nos elementos")
            elif mc == "elmtpresloads":
                _bound(src, m["eloadp"], numel, 7, "cargas nos elementos")
            else:
                _initial(src, m["u0"], nnode, 3)
        elif mc == "loads":
            _rload(src, m["load"], m["fload"])
        elif mc == "hydrostatic":
            if not have_x:
                raise DeckError("hydrostatic before coordinates")
            _hydrostatic(src.path(w[1]) if len(w) > 1 else "", m["x"], m["u0"], st["gravity_mod"])
        elif mc == "initialdisp":
            _skip_block(src)  # read and dropped by the reference (Rdata.f:1600-1607)
        elif mc == "insert":
            if len(w) < 2:
                raise DeckError("insert without a file name")
            src.push(src.path(w[1]))
        elif mc == "return":
            src.pop()
        elif mc in UNSUPPORTED_MESH:
            raise DeckError(f"mesh macro {mc!r} is outside the mef Biot u-p path")
    if not have_el or not have_x:
        raise DeckError("mesh block without hexa8 elements or coordinates")
    mats = m["ix"][:, NEN]
    if numel and (mats.min() < 1 or mats.max() > numat):
        raise DeckError("element material id out of range")
    # boundc (Rdata.f:1650, 2300): restrained dofs start at their prescribed
    # values, which are zero without nodalforces; then u = u0
    m["u0"][m["idv"] > 0] = 0.0
    return m


def _parameters(src):
    """The mesh parameter line (parameters, Rdata.f:1863-1943)."""
    w = src.next()
    if w is None:
        raise DeckError("mesh without parameters")
    par = {}
    keys = ("nnode", "numel", "numat", "maxno", "ndf", "dim")
    i = 0
    while i < len(w):
        if w[i][:8] in keys:
            if i + 1 >= len(w):
                raise DeckError("*** Erro in reading the control variables !")
            par[w[i][:8]] = inum(w[i + 1])
            i += 2
        else:
            i += 1
    for k in keys:
        if k not in par:
            raise DeckError(f"Missing macro: {k}")
    return par


def _materials(src, m, numat):
    """``materials``: ``ma type e1 e2 ...`` per line until ``end`` (mate,
    Rdata.f:1132-1182)."""
    while True:
        w = src.next()
        if w is None:
            raise DeckError("*** Erro na leitura dos materiais !")
        if not w:
            continue
        if w[0] == "end":
            return
        ma = inum(w[0])
        if not 1 <= ma <= numat or len(w) < 2:
            raise DeckError("*** Erro na leitura dos materiais !")
        _element_type(m, inum(w[1]))
        vals = w[2:]
        if len(vals) > PROP:
            raise DeckError("*** Erro na leitura dos materiais !")
        m["ie"][ma - 1] = inum(w[1])
        for i, v in enumerate(vals):
            m["e"][ma - 1, i] = fnum(v)


def _fmaterials(src, m, numat):
    """``fmaterials``: ``ma file`` per line; each file holds ``key value``
    pairs until ``end`` (fmate/file_prop, Rdata.f:1184-1240, 2683-2850)."""
    keys = ("modE", "poisson", "permeability", "mbiot", "cbiot", "density", "fdensity",
            "porosity", "l_plastic", "k_plastic", "mcs", "pc0")
    while True:
        w = src.next()
        if w is None:
            raise DeckError("*** Erro na leitura dos materiais !")
        if not w:
            continue
        if w[0] == "end":
            return
        ma = inum(w[0])
        if not 1 <= ma <= numat or len(w) < 2:
            raise DeckError("*** Erro na leitura dos materiais !")
        words = _file_words(src.path(w[1]))
        i, etype = 0, None
        while i + 1 < len(words) and words[i] != "end":
            if words[i] in keys:
                m["e"][ma - 1, keys.index(words[i])] = fnum(words[i + 1])
            elif words[i] == "elType":
                etype = inum(words[i + 1])
            i += 2
        if etype is None:
            raise DeckError(f"material {ma}: elType missing")
        _element_type(m, etype)
        m["ie"][ma - 1] = etype


def _element_type(m, ty):
    if ty not in ELEMENT_TYPES:
        raise DeckError(f"element type {ty}: this path is the hexa20/8 Biot element "
                        "(17, or 18 with properties at the integration points)")
    m["etype"] = ty


def _elconn(src, ix, numel, nenv):
    """``hexa8``: exactly numel lines ``k v1..v8 mat``, then ``end`` (elconn,
    Rdata.f:1038-1088)."""
    for _ in range(numel):
        w = _data_line(src, "*** Erro na leitura dos elementos !")
        if len(w) < nenv + 2:
            raise DeckError("*** Erro na leitura dos elementos !")
        k = inum(w[0])
        if not 1 <= k <= numel:
            raise DeckError("*** Erro na leitura dos elementos !")
        ix[k - 1, :nenv] = [inum(v) for v in w[1:nenv + 1]]
        ix[k - 1, NEN] = inum(w[nenv + 1])
    w = _data_line(src, "*** Erro na leitura dos elementos !")
    if w[0] != "end":
        raise DeckError("*** Erro na leitura dos elementos !")


def _coord(src, x, nnode):
    """``coordinates``: exactly nnode lines ``k x y z``, then ``end`` (coord,
    Rdata.f:1090-1130)."""
    for _ in range(nnode):
        w = _data_line(src, "*** Erro na leitura das coordenadas !")
        if len(w) < 4:
            raise DeckError("*** Erro na leitura das coordenadas !")
        k = inum(w[0])
        if not 1 <= k <= nnode:
            raise DeckError("*** Erro na leitura das coordenadas !")
        x[k - 1] = [fnum(v) for v in w[1:4]]
    w = _data_line(src, "*** Erro na leitura das coordenadas !")
    if w[0] != "end":
        raise DeckError("*** Erro na leitura das coordenadas !")


def _bound(src, arr, n, ndf, what):
    """``j v1 .. vk`` lines (k <= ndf, the rest stay zero) until ``end``
    (bound, Rdata.f:1408-1460)."""
    while True:
        w = _data_line(src, f"*** Erro na leitura das {what} !")
        if w[0] == "end":
            return
        j = inum(w[0])
        if not 1 <= j <= n or len(w) - 1 > ndf:
            raise DeckError(f"*** Erro na leitura das {what} ! {j}")
        for i, v in enumerate(w[1:]):
            arr[j - 1, i] = inum(v)


def _initial(src, u0, n, col):
    """``initialpres``: ``k p`` lines until ``end`` into u0(4, k)
    (init_poro_mec with n1 = n2 = ndf, Rdata.f:1267-1298)."""
    while True:
        w = _data_line(src, "*** Erro na leitura das forcas nodais !")
        if w[0] == "end":
            return
        k = inum(w[0])
        if not 1 <= k <= n or len(w) < 2:
            raise DeckError("*** Erro na leitura das forcas nodais !")
        u0[k - 1, col] = fnum(w[1])


def _hydrostatic(path, x, u0, gmod):
    """``hydrostatic <file>``: density, p0, datum h, scale; every vertex gets
    p = (p0 + density * |g| * (h - z)) * scale (init_hydrostatic_pres,
    Rdata.f:1301-1345)."""
    try:
        words = _file_words(path)
    except OSError:
        raise DeckError(f"File {path} not found !") from None
    if len(words) < 4:
        raise DeckError("*** Erro na leitura da funcao init_hydrostatic_pres !")
    dens, p0, h, scale = (fnum(v) for v in words[:4])
    u0[:, 3] = (p0 + dens * gmod * (h - x[:, 2])) * scale


def _skip_block(src):
    while True:
        w = src.next()
        if w is None or (w and w[0] == "end"):
            return


def _data_line(src, msg):
    while True:
        w = src.next()
        if w is None:
            raise DeckError(msg)
        if w:
            return w


def _rload(src, load, fload):
    """``loads`` (rload, Rdata.f:1487-1690).  ``fload`` is indexed
    [nc, term, parameter] (Fortran fload(parameter, term, nc))."""
    while True:
        w = _data_line(src, "*** Load reading error !")
        if w[0] == "end":
            return
        nc = inum(w[0])
        if not 1 <= nc <= NUMLOAD or len(w) < 2:
            raise DeckError(f"*** Load reading error ! {nc}")
        ty = inum(w[1])
        load[nc - 1, 0] = ty
        load[nc - 1, 1] = 1
        a = w[2:]
        i = nc - 1
        if ty == 1:
            fload[i, 0, 0] = fnum(a[0])
        elif ty in (2, 3):
            fload[i, 0, 0], fload[i, 0, 1] = fnum(a[0]), fnum(a[1])
        elif ty in (4, 5):
            raise DeckError(f"load type {ty}: rload reads no values for it")
        elif ty == 6:
            # n terms of (amplitude, period, phase), one term per line
            n = inum(a[0])
            if not 0 <= n <= NTERMS:
                raise DeckError(f"*** Load reading error ! {nc}")
            load[i, 1] = n
            for k in range(n):
                t = _data_line(src, "*** Load reading error !")
                for p in range(3):
                    fload[i, k, p] = fnum(t[p])
        elif ty == 7:
            for p in range(3):
                fload[i, 0, p] = fnum(a[p])
            n = inum(a[3])
            if not 1 <= n <= NPAR:
                raise DeckError(f"*** Load reading error ! {nc}")
            load[i, 1] = n
            for k in range(n):
                t = _data_line(src, "*** Load reading error !")
                fload[i, 1, k], fload[i, 2, k] = fnum(t[0]), fnum(t[1])
        elif ty == 8:
            n = inum(a[0])
            if not 1 <= n <= NPAR:
                raise DeckError(f"*** Load reading error ! {nc}")
            load[i, 1] = n
            for k in range(n):
                t = _data_line(src, "*** Load reading error !")
                fload[i, 0, k], fload[i, 1, k] = fnum(t[0]), fnum(t[1])
        elif 39 <= ty <= 43:
            if len(a) == 1:  # <file>: nterms ntimes, then ntimes rows of nterms values
                try:
                    words = _file_words(src.path(a[0]))
                except OSError:
                    raise DeckError(f"File {a[0]} not found !") from None
                nterm, ntime = inum(words[0]), inum(words[1])
                if not 1 <= nterm <= NTERMS or not 1 <= ntime <= NPAR:
                    raise DeckError(f"*** Load reading error ! {nc}")
                if len(words) < 2 + nterm * ntime:
                    raise DeckError(f"*** Load reading error ! {nc}")
                vals = [fnum(v) for v in words[2:2 + nterm * ntime]]
                load[i, 1], load[i, 2] = nterm, ntime
                for ll in range(ntime):
                    for k in range(nterm):
                        fload[i, k, ll] = vals[ll * nterm + k]
            else:  # earlier inline form: nc type n v1 .. vn, constant in time
                n = inum(a[0])
                if not 1 <= n <= NTERMS - 1 or len(a) < 1 + n:
                    raise DeckError(f"*** Load reading error ! {nc}")
                load[i, 1], load[i, 2] = n + 1, 2
                for ll, t in enumerate((0.0, 1.0e30)):
                    fload[i, 0, ll] = t
                    for k in range(n):
                        fload[i, 1 + k, ll] = fnum(a[1 + k])
        else:
            raise DeckError(f"load type {ty} is outside this path")


# -- writer ----------------------------------------------------------------------------
# columns per time row of a file load (time + values), by load type (tload,
# Pload.f:330-370: 39 / 41 / 43 one value, 40 a vector, 42 four parameters)
FILE_LOAD_TERMS = {39: 2, 40: 4, 41: 2, 42: 5, 43: 2}
SOLVER_FILE_NAMES = {"pcg": "cg", "gmres": "gmres", "bicgstab": "bicgs", "minres": "minres",
                     "sqmr": "sqmr"}
PRECOND_FILE_NAMES = {1: "none", 2: "diag", 5: "diagm"}


def _f(v) -> str:
    return repr(float(v))


def truncate(inp: MefInput, nsteps: int) -> MefInput:
    """The same input with the macro program cut after its nsteps-th
    ``solv`` (and the print macros that follow it before the next one)."""
    out, n = [], 0
    for op in inp.program:
        if op[0] == "solv":
            if n == nsteps:
                break
            n += 1
        out.append(op)
    return dataclasses.replace(inp, program=out)


def write_deck(path, deck: Deck, cfg: SolverConfig | None = None, u0=None, reorder=True,
               pnodes=(1,), nsteps=None, element_type=17, vprop=None):
    """Writes ``deck`` as an input deck of the reference in its current
    grammar: the main file ``path`` and the files it names beside it
    (``<stem>_config.dat``, ``_solver.dat``, ``_gravity.dat``, ``_node.dat``,
    ``_load<nc>.dat``, ``_conseq.dat``).  ``element_type`` is one type for
    every material or a list per material; ``vprop`` the conseq flags
    (vprop, kozeny_carman, density, hashin_shtrikman).  The program is
    ``dt``, ``pnup``, then ``nsteps`` (default ``deck.nsteps``) times
    ``solv`` + ``pnup``.  A deck with ``nen = 8`` is written with ``maxno
    8`` and element type 7 (Q1/Q1).  Returns the path."""
    cfg = cfg or SolverConfig()
    nsteps = deck.nsteps if nsteps is None else nsteps
    path = os.fspath(path)
    base, stem = os.path.dirname(path), os.path.splitext(os.path.basename(path))[0]

    def side(suffix, text):
        name = f"{stem}_{suffix}.dat"
        with open(os.path.join(base, name), "w") as f:
            f.write(text)
        return name

    numel, nnode = deck.numel, deck.nnodev
    numat = deck.e.shape[0]
    L = []
    L.append("config " + side("config", f"reord {'true' if reorder else 'false'}\nend\n"))
    L.append("solver " + side("solver", (
        f"name {SOLVER_FILE_NAMES[cfg.solver]}\nit {cfg.maxit}\ntol {_f(cfg.tol)}\n"
        f"precond {PRECOND_FILE_NAMES[cfg.precond]}\nnkrylov {cfg.ngram}\nend\n")))
    g = deck.gravity
    L.append("gravity " + side("gravity", f"{_f(g[0])} {_f(g[1])} {_f(g[2])} "
                                          f"{_f(deck.gravity_mod)}\n"))
    L.append(f"nl {cfg.maxnlit} {_f(cfg.nltol)} {cfg.stop_crit}")
    if vprop is not None and any(vprop):
        names = ("vprop", "kozeny_carman", "density", "hashin_shtrikman")
        L.append("conseq " + side("conseq", "".join(f"{n}\n" for n, f in zip(names, vprop) if f)
                                  + "end\n"))
    L.append("mesh")
    L.append(f"nnode {nnode} numel {numel} numat {numat} maxno {deck.nen} ndf {NDF} dim 3")
    L.append("materials")
    types = np.broadcast_to(np.asarray(7 if deck.nen == 8 else element_type), (numat,))
    for ma in range(numat):
        row = deck.e[ma]
        last = max(7, int(np.nonzero(row)[0].max()) + 1 if row.any() else 7)
        L.append(f"{ma + 1} {int(types[ma])} " + " ".join(_f(v) for v in row[:last]))
    L.append("end materials")
    L.append("coordinates")
    L.extend(f"{k + 1} {_f(x[0])} {_f(x[1])} {_f(x[2])}" for k, x in enumerate(deck.x))
    L.append("end coordinates")
    L.append("hexa8")
    L.extend(f"{k + 1} " + " ".join(str(int(v)) for v in r[:8]) + f" {int(r[NEN])}"
             for k, r in enumerate(deck.ix))
    L.append("end hexa8")
    for name, arr in (("constrainpmec", deck.idv), ("elmtloads", deck.eload),
                      ("elmtpresloads", deck.eloadp)):
        L.append(name)
        L.extend(f"{k + 1} " + " ".join(str(int(v)) for v in r)
                 for k, r in enumerate(arr) if np.any(r))
        L.append(f"end {name}")
    L.append("loads")
    for i in range(NUMLOAD):
        ty = int(deck.load[i, 0])
        if ty == 0:
            continue
        if ty in FILE_LOAD_TERMS:
            nterm, ntime = FILE_LOAD_TERMS[ty], int(deck.load[i, 2])
            rows = "\n".join(" ".join(_f(deck.fload[i, k, ll]) for k in range(nterm))
                             for ll in range(ntime))
            L.append(f"{i + 1} {ty} " + side(f"load{i + 1}", f"{nterm} {ntime}\n{rows}\n"))
        elif ty == 1:
            L.append(f"{i + 1} 1 {_f(deck.fload[i, 0, 0])}")
        elif ty in (2, 3):
            L.append(f"{i + 1} {ty} {_f(deck.fload[i, 0, 0])} {_f(deck.fload[i, 0, 1])}")
        else:
            raise DeckError(f"load type {ty}: no writer")
    L.append("end loads")
    if u0 is not None and np.any(np.asarray(u0)[:, 3]):
        L.append("initialpres")
        L.extend(f"{k + 1} {_f(p)}" for k, p in enumerate(np.asarray(u0)[:, 3]))
        L.append("end initialpres")
    L.append("end mesh")
    L.append("setpnode " + side("node", f"{len(pnodes)}\n" + "".join(f"{n}\n" for n in pnodes)))
    L.append(f"dt {_f(deck.dt)}")
    L.append("pnup")
    L.append(f"loop {nsteps}")
    L.append("solv")
    L.append("pnup")
    L.append("next")
    L.append("stop")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")
    return path
