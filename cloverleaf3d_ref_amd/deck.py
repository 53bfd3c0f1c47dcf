"""Input decks: the reference's ``*clover ... *endclover`` format.

Restates the deck reader of the reference (read_input.f90:24-305 with the
line rules of parse.f90:143-247): text between ``*clover`` and ``*endclover``
is lower-cased, ``!`` and ``;`` start comments, ``=`` and ``,`` separate words
like blanks, and keywords are followed by their value.  Defaults are those of
read_input.f90:38-86.  After parsing, every state other than the background
is shrunk by a hundredth of a cell on each side (read_input.f90:293-303), so
that state boundaries never fall exactly on a cell face.
"""
from __future__ import annotations

import dataclasses
import re
from typing import List, Optional

G_RECT, G_CIRC, G_POINT = 1, 2, 3  # data.f90:90-92
G_IBIG = 640000                    # data.f90:28 (default end_step)


@dataclasses.dataclass
class State:
    density: float = 0.0
    energy: float = 0.0
    xvel: float = 0.0
    yvel: float = 0.0
    zvel: float = 0.0
    xmin: float = 0.0
    xmax: float = 0.0
    ymin: float = 0.0
    ymax: float = 0.0
    zmin: float = 0.0
    zmax: float = 0.0
    radius: float = 0.0
    geometry: int = 0
    defined: bool = False


@dataclasses.dataclass
class Deck:
    x_cells: int = 10
    y_cells: int = 10
    z_cells: int = 10
    xmin: float = 0.0
    ymin: float = 0.0
    zmin: float = 0.0
    xmax: float = 100.0
    ymax: float = 100.0
    zmax: float = 100.0
    dtinit: float = 0.1
    dtmax: float = 1.0
    dtmin: float = 0.0000001
    dtrise: float = 1.05
    dtc_safe: float = 0.75
    dtu_safe: float = 0.5
    dtv_safe: float = 0.5
    dtw_safe: float = 0.5
    dtdiv_safe: float = 0.5
    end_time: float = 10.0
    end_step: int = G_IBIG
    visit_frequency: int = 0
    summary_frequency: int = 10
    tiles_per_chunk: int = 1
    tile_dims: int = 3
    profiler_on: bool = False
    test_problem: int = 0
    states: List[State] = dataclasses.field(default_factory=list)

    @property
    def cells(self) -> int:
        return self.x_cells * self.y_cells * self.z_cells


class DeckError(ValueError):
    """A deck the reference would reject (read_input.f90:108, 213)."""


def _getrval(word: str) -> float:
    # parse.f90:332-354 appends '.0' to integers, then reads a REAL(8)
    return float(word)


def _getival(word: str) -> int:
    return int(word)


def _section_words(text: str) -> List[List[str]]:
    lines: List[List[str]] = []
    inside = False
    for raw in text.splitlines():
        line = "".join(c if 32 <= ord(c) <= 128 else " " for c in raw).strip()
        for mark in ("!", ";"):
            cut = line.find(mark)
            if cut >= 0:
                line = line[:cut].strip()
        line = line.lower()
        line = re.sub(r"[=,]", " ", line)
        if not line:
            continue
        if line.startswith("*"):
            tag = line.split()[0]
            if tag == "*clover":
                inside = True
            elif tag == "*endclover":
                inside = False
            continue
        if inside:
            lines.append(line.split())
    return lines


def parse_deck(text: str) -> Deck:
    """Parse deck text into a :class:`Deck` (read_input.f90:24-305)."""
    lines = _section_words(text)
    state_max = 0
    for words in lines:
        if "state" in words:
            i = words.index("state")
            if i + 1 < len(words):
                state_max = max(state_max, _getival(words[i + 1]))
    if state_max < 1:
        raise DeckError("read_input: No states defined.")
    deck = Deck()
    deck.states = [State() for _ in range(state_max)]

    for words in lines:
        i = 0
        while i < len(words):
            w = words[i]
            nxt = words[i + 1] if i + 1 < len(words) else None

            def val() -> str:
                if nxt is None:
                    raise DeckError(f"read_input: missing value for {w}")
                return nxt

            if w == "state":
                sid = _getival(val())
                st = deck.states[sid - 1]
                if st.defined:
                    raise DeckError("read_input: State defined twice.")
                st.defined = True
                j = i + 2
                while j < len(words):
                    key = words[j]
                    arg = words[j + 1] if j + 1 < len(words) else None
                    if key in ("xvel", "yvel", "zvel", "xmin", "ymin", "zmin", "xmax", "ymax", "zmax",
                               "radius", "density", "energy"):
                        if arg is None:
                            raise DeckError(f"read_input: missing value for state {key}")
                        setattr(st, key, _getrval(arg))
                        j += 2
                    elif key == "geometry":
                        st.geometry = {"cuboid": G_RECT, "sphere": G_CIRC, "point": G_POINT}.get(arg or "", 0)
                        j += 2
                    else:
                        j += 1
                break
            if w in ("initial_timestep", "max_timestep", "timestep_rise", "end_time", "xmin", "xmax",
                     "ymin", "ymax", "zmin", "zmax"):
                attr = {"initial_timestep": "dtinit", "max_timestep": "dtmax",
                        "timestep_rise": "dtrise"}.get(w, w)
                setattr(deck, attr, _getrval(val()))
                i += 2
            elif w in ("end_step", "x_cells", "y_cells", "z_cells", "visit_frequency",
                       "summary_frequency", "tiles_per_chunk", "test_problem"):
                setattr(deck, w, _getival(val()))
                i += 2
            elif w == "tiles_per_problem":
                deck.tiles_per_chunk = _getival(val())  # one chunk: tiles_per_problem/1
                i += 2
            elif w in ("1d_tile_decomposition", "2d_tile_decomposition", "3d_tile_decomposition"):
                deck.tile_dims = int(w[0])
                i += 1
            elif w == "profiler_on":
                deck.profiler_on = True
                i += 1
            else:
                i += 1

    # read_input.f90:293-303
    dx = (deck.xmax - deck.xmin) / float(deck.x_cells)
    dy = (deck.ymax - deck.ymin) / float(deck.y_cells)
    dz = (deck.zmax - deck.zmin) / float(deck.z_cells)
    for st in deck.states[1:]:
        st.xmin = st.xmin + (dx / 100.0)
        st.ymin = st.ymin + (dy / 100.0)
        st.zmin = st.zmin + (dz / 100.0)
        st.xmax = st.xmax - (dx / 100.0)
        st.ymax = st.ymax - (dy / 100.0)
        st.zmax = st.zmax - (dz / 100.0)
    return deck


# The deck that initialise.f90:77-93 writes when clover.in is missing; its
# final kinetic energy is the reference's only known answer
# (field_summary.f90:119-136).
DEFAULT_DECK = """*clover
 state 1 density=0.2 energy=1.0
 state 2 density=1.0 energy=2.5 geometry=cuboid xmin=0.0 xmax=5.0 ymin=0.0 ymax=2.0 zmin=0.0 zmax=2.0
 x_cells=10
 y_cells=2
 z_cells=2
 xmin=0.0
 ymin=0.0
 zmin=0.0
 xmax=10.0
 ymax=2.0
 zmax=2.0
 initial_timestep=0.04
 max_timestep=0.04
 end_step=75
 test_problem 1
*endclover
"""
TEST_PROBLEM_1_KE = 3.64560737191257  # field_summary.f90:123


def cuboid_deck(cells: int, end_step: int, summary_frequency: int, tiles: int = 1,
                ny: Optional[int] = None, nz: Optional[int] = None) -> str:
    """The two-state cuboid deck of the benchmark configurations: [0,10]^3,
    background rho 0.2 e 1.0, cuboid x in [0,5], y,z in [0,2] at rho 1.0
    e 2.5, dt0 0.04, rise 1.5, end_time large so end_step terminates."""
    ny = cells if ny is None else ny
    nz = cells if nz is None else nz
    text = f"""*clover
 state 1 density=0.2 energy=1.0
 state 2 density=1.0 energy=2.5 geometry=cuboid xmin=0.0 xmax=5.0 ymin=0.0 ymax=2.0 zmin=0.0 zmax=2.0
 x_cells={cells}
 y_cells={ny}
 z_cells={nz}
 xmin=0.0
 ymin=0.0
 zmin=0.0
 xmax=10.0
 ymax=10.0
 zmax=10.0
 initial_timestep=0.04
 timestep_rise=1.5
 end_time=1000.0
 end_step={end_step}
 summary_frequency={summary_frequency}
 tiles_per_chunk {tiles}
"""
    if tiles > 1:
        text += " 3d_tile_decomposition\n"
    return text + "*endclover\n"


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014): the ensemble generator's PRNG."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.s = seed & self.MASK

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        """Uniform on [lo, hi) from the top 53 bits."""
        return lo + (hi - lo) * ((self.next() >> 11) * 2.0 ** -53)


ENSEMBLE_SEED = 0x434C4F5645523344  # fixed seed of the batched workload
ENSEMBLE_EDGES = tuple(range(16, 97, 8))  # 16, 24, ..., 96


def ensemble_deck_texts(n: int, seed: int = ENSEMBLE_SEED, end_step: int = 200,
                        edges=ENSEMBLE_EDGES) -> List[str]:
    """Input decks of the batched ensemble: per deck an edge drawn uniformly
    from `edges` on every axis, background density U[0.1,0.5] and energy
    U[0.5,1.5], one cuboid of density U[0.5,2.0] and energy U[1.5,5.0]
    covering U[10%,40%] of each axis at a uniform position, on the domain
    [0,10]^3 with the cuboid deck's time stepping (dt0 0.04, rise 1.5)."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(n):
        cells = [edges[int(rng.uniform() * len(edges))] for _ in range(3)]
        rho_b, e_b = rng.uniform(0.1, 0.5), rng.uniform(0.5, 1.5)
        rho_c, e_c = rng.uniform(0.5, 2.0), rng.uniform(1.5, 5.0)
        box = []
        for _ in range(3):
            length = 10.0 * rng.uniform(0.1, 0.4)
            lo = rng.uniform(0.0, 10.0 - length)
            box += [lo, lo + length]
        out.append(f"""*clover
 state 1 density={rho_b!r} energy={e_b!r}
 state 2 density={rho_c!r} energy={e_c!r} geometry=cuboid xmin={box[0]!r} xmax={box[1]!r} ymin={box[2]!r} ymax={box[3]!r} zmin={box[4]!r} zmax={box[5]!r}
 x_cells={cells[0]}
 y_cells={cells[1]}
 z_cells={cells[2]}
 xmin=0.0
 ymin=0.0
 zmin=0.0
 xmax=10.0
 ymax=10.0
 zmax=10.0
 initial_timestep=0.04
 timestep_rise=1.5
 end_time=1000.0
 end_step={end_step}
 summary_frequency=0
 tiles_per_chunk 1
*endclover
""")
    return out


def ensemble_decks(n: int, seed: int = ENSEMBLE_SEED, end_step: int = 200, edges=ENSEMBLE_EDGES) -> List[Deck]:
    return [parse_deck(t) for t in ensemble_deck_texts(n, seed, end_step, edges)]
