"""Seeded synthetic inputs for the benchmark configurations (vectorized numpy; the mixed batch uses torch).

Each generator returns host tapes (``Strs``) plus the parameters of the engine that scores them. Seeds are
``0x535A0000 + config_index`` (numpy's PCG64, or torch's generator for the mixed batch). Choices the configuration text leaves open are fixed here:

* lognormal lengths use sigma = 0.5 around the stated median;
* the two sides of a pair draw their lengths independently;
* UTF-8 strings draw codepoints uniformly from ASCII printable (U+0020-007E), Cyrillic (U+0400-04FF), CJK
  (U+4E00-9FFF) and emoji (U+1F300-1F5FF) ranges, one range per codepoint chosen uniformly, and stop before the
  drawn byte length would be exceeded;
* the NW matrix has 21 classes: class 0 for every byte outside the 20 amino-acid letters, classes 1-20 for them,
  symmetric, diagonal uniform in [4, 11], off-diagonal uniform in [-4, 3];
* in the mixed batch the 32 extra 65,536-byte pairs follow the same kind split (16 byte Levenshtein, 8 UTF-8,
  8 NW).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

AMINO = np.frombuffer(b"ACDEFGHIKLMNPQRSTVWY", dtype=np.uint8)
DNA = np.frombuffer(b"ACGT", dtype=np.uint8)
SEED_BASE = 0x535A0000
UTF8_RANGES = [(0x20, 0x7E), (0x400, 0x4FF), (0x4E00, 0x9FFF), (0x1F300, 0x1F5FF)]


@dataclass
class Segment:
    """One homogeneous pair list: ``kind`` is levenshtein / levenshtein_utf8 / needleman_wunsch / smith_waterman."""

    kind: str
    firsts_data: np.ndarray
    firsts_offsets: np.ndarray
    seconds_data: np.ndarray
    seconds_offsets: np.ndarray
    match: int = 0
    mismatch: int = 1
    open: int = 1
    extend: int = 1
    byte_to_class: Optional[np.ndarray] = None
    class_costs: Optional[np.ndarray] = None

    @property
    def count(self) -> int:
        return len(self.firsts_offsets) - 1

    def host(self) -> "Segment":
        """The same segment with numpy arrays (copies device tensors to the host)."""
        def np_(x):
            return x.cpu().numpy() if hasattr(x, "cpu") else x

        return Segment(self.kind, np_(self.firsts_data).astype(np.uint8), np_(self.firsts_offsets).astype(np.uint64),
                       np_(self.seconds_data).astype(np.uint8), np_(self.seconds_offsets).astype(np.uint64),
                       self.match, self.mismatch, self.open, self.extend, self.byte_to_class, self.class_costs)

    def lengths(self, which: str = "firsts") -> np.ndarray:
        offsets = self.firsts_offsets if which == "firsts" else self.seconds_offsets
        data = self.firsts_data if which == "firsts" else self.seconds_data
        if hasattr(offsets, "cpu"):
            import torch

            offsets = offsets.to(torch.int64)
            if self.kind != "levenshtein_utf8":
                return (offsets[1:] - offsets[:-1]).cpu().numpy()
            is_lead = ((data & 0xC0) != 0x80).to(torch.int64)
            cumulative = torch.cat([torch.zeros(1, dtype=torch.int64, device=data.device), torch.cumsum(is_lead, 0)])
            return (cumulative[offsets[1:]] - cumulative[offsets[:-1]]).cpu().numpy()
        if self.kind != "levenshtein_utf8":
            return np.diff(offsets).astype(np.int64)
        # Rune count = bytes that are not UTF-8 continuation bytes (inputs here are valid UTF-8).
        is_lead = (data & 0xC0) != 0x80
        cumulative = np.concatenate([[0], np.cumsum(is_lead, dtype=np.int64)])
        return cumulative[offsets[1:].astype(np.int64)] - cumulative[offsets[:-1].astype(np.int64)]

    def cells(self) -> int:
        return int(np.dot(self.lengths("firsts").astype(np.int64), self.lengths("seconds").astype(np.int64)))

    def ops_per_cell(self) -> int:
        """Fixed VALU ops per cell update of the recurrence family (linear 5, affine global 9, affine local 10)."""
        affine = self.open != self.extend
        if not affine:
            return 5
        return 10 if self.kind == "smith_waterman" else 9

    def subset(self, indices: np.ndarray) -> "Segment":
        a_data, a_off = _gather(self.firsts_data, self.firsts_offsets, indices)
        b_data, b_off = _gather(self.seconds_data, self.seconds_offsets, indices)
        return Segment(self.kind, a_data, a_off, b_data, b_off, self.match, self.mismatch, self.open, self.extend,
                       self.byte_to_class, self.class_costs)


@dataclass
class Workload:
    name: str
    segments: List[Segment] = field(default_factory=list)

    @property
    def pairs(self) -> int:
        return sum(s.count for s in self.segments)

    def cells(self) -> int:
        return sum(s.cells() for s in self.segments)

    def weighted_ops_per_cell(self) -> float:
        cells = [s.cells() for s in self.segments]
        total = sum(cells)
        return sum(c * s.ops_per_cell() for c, s in zip(cells, self.segments)) / max(total, 1)


def _gather(data: np.ndarray, offsets: np.ndarray, indices: np.ndarray):
    begins = offsets[:-1][indices].astype(np.int64)
    ends = offsets[1:][indices].astype(np.int64)
    lengths = ends - begins
    new_offsets = np.zeros(len(indices) + 1, dtype=np.uint64)
    new_offsets[1:] = np.cumsum(lengths)
    total = int(new_offsets[-1])
    if total == 0:
        return np.zeros(1, dtype=np.uint8), new_offsets
    starts = np.repeat(begins - new_offsets[:-1].astype(np.int64), lengths)
    positions = np.arange(total, dtype=np.int64) + starts
    out = np.empty(total + 1, dtype=np.uint8)
    out[:total] = data[positions]
    out[total] = 0
    return out, new_offsets


def _offsets(lengths: np.ndarray) -> np.ndarray:
    offsets = np.zeros(len(lengths) + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum(lengths.astype(np.uint64))
    return offsets


def _random_tape(rng, alphabet: np.ndarray, lengths: np.ndarray):
    offsets = _offsets(lengths)
    total = int(offsets[-1])
    raw = np.frombuffer(rng.bytes(total + 1), dtype=np.uint8)
    if len(alphabet) & (len(alphabet) - 1) == 0:
        data = alphabet[raw & (len(alphabet) - 1)]
    else:
        data = alphabet[(raw.astype(np.uint16) * len(alphabet)) >> 8]
    return np.ascontiguousarray(data, dtype=np.uint8), offsets


def protein_matrix(rng):
    """21-class symmetric matrix: diagonal U[4, 11], off-diagonal U[-4, 3]; classes 1-20 = the amino letters."""
    byte_to_class = np.zeros(256, dtype=np.uint8)
    byte_to_class[AMINO] = np.arange(1, 21, dtype=np.uint8)
    costs = np.zeros((32, 32), dtype=np.int8)
    upper = rng.integers(-4, 4, size=(21, 21))
    block = np.triu(upper, 1)
    block = block + block.T
    np.fill_diagonal(block, rng.integers(4, 12, size=21))
    costs[:21, :21] = block
    return byte_to_class, np.ascontiguousarray(costs)


def dense_matrix_256(byte_to_class: np.ndarray, class_costs: np.ndarray) -> np.ndarray:
    """The 256x256 int8 view of a class-compressed matrix: ``M[a][b] = C[class(a)][class(b)]``."""
    return class_costs[byte_to_class[:, None], byte_to_class[None, :]]


def levenshtein_c0(count: int = 1024) -> Workload:
    """1024 ACGT pairs, lengths uniform in [64, 256], unit costs."""
    rng = np.random.default_rng(SEED_BASE + 0)
    a = _random_tape(rng, DNA, rng.integers(64, 257, size=count))
    b = _random_tape(rng, DNA, rng.integers(64, 257, size=count))
    return Workload("levenshtein_1024_acgt_64_256", [Segment("levenshtein", *a, *b)])


def needleman_wunsch_c1(count: int = 16384) -> Workload:
    """Affine NW (-11, -1), 20-letter alphabet, lengths lognormal median 350 (sigma 0.5) clipped to [50, 2000]."""
    rng = np.random.default_rng(SEED_BASE + 1)
    b2c, costs = protein_matrix(rng)

    def lengths():
        return np.clip(np.exp(rng.normal(np.log(350), 0.5, size=count)).astype(np.int64), 50, 2000)

    a = _random_tape(rng, AMINO, lengths())
    b = _random_tape(rng, AMINO, lengths())
    return Workload("needleman_wunsch_affine_16384_protein",
                    [Segment("needleman_wunsch", *a, *b, open=-11, extend=-1, byte_to_class=b2c, class_costs=costs)])


def smith_waterman_c2(count: int = 100_000) -> Workload:
    """SW linear -2, +2/-3; query 150 ACGT; target U[150, 2000] embedding a mutated query copy in random flanks."""
    rng = np.random.default_rng(SEED_BASE + 2)
    queries = DNA[rng.integers(0, 4, size=(count, 150))]
    targets = []
    target_lengths = rng.integers(150, 2001, size=count)
    for i in range(count):
        query = queries[i]
        copy = []
        position = 0
        while position < 150:
            roll = rng.random()
            if roll < 0.01:  # indel of length 1-3
                size = int(rng.integers(1, 4))
                if rng.random() < 0.5:
                    copy.extend(DNA[rng.integers(0, 4, size=size)])
                else:
                    position += size
                continue
            symbol = query[position]
            if roll < 0.06:
                symbol = DNA[(np.searchsorted(DNA, symbol) + rng.integers(1, 4)) % 4]
            copy.append(symbol)
            position += 1
        copy = np.array(copy, dtype=np.uint8)
        length = max(int(target_lengths[i]), len(copy))
        target = DNA[rng.integers(0, 4, size=length)]
        offset = int(rng.integers(0, length - len(copy) + 1))
        target[offset:offset + len(copy)] = copy
        targets.append(target.tobytes())
    q_offsets = np.arange(count + 1, dtype=np.uint64) * 150
    q_data = np.concatenate([queries.reshape(-1), [0]]).astype(np.uint8)
    t_lengths = np.array([len(t) for t in targets], dtype=np.int64)
    t_data = np.frombuffer(b"".join(targets) + b"\0", dtype=np.uint8)
    b2c = np.zeros(256, dtype=np.uint8)
    b2c[DNA] = np.arange(1, 5, dtype=np.uint8)
    costs = np.full((32, 32), -3, dtype=np.int8)
    np.fill_diagonal(costs, 2)
    return Workload("smith_waterman_linear_100000_dna",
                    [Segment("smith_waterman", q_data, q_offsets, t_data, _offsets(t_lengths), open=-2, extend=-2,
                             byte_to_class=b2c, class_costs=costs)])


def _torch_offsets(lengths):
    import torch

    offsets = torch.zeros(len(lengths) + 1, dtype=torch.int64, device=lengths.device)
    offsets[1:] = torch.cumsum(lengths, 0)
    return offsets


def _torch_random_tape(generator, alphabet: np.ndarray, lengths):
    import torch

    device = lengths.device
    offsets = _torch_offsets(lengths)
    total = int(offsets[-1])
    table = torch.from_numpy(alphabet.copy()).to(device)
    picks = torch.randint(0, len(alphabet), (total + 1,), generator=generator, device=device, dtype=torch.int64)
    data = table[picks]
    data[total] = 0
    return data, offsets


def _torch_utf8_tape(generator, budgets):
    """Torch twin of the UTF-8 generator above: same distribution, any device."""
    import torch

    device = budgets.device
    draws = budgets // 2 + 16
    total_draws = int(draws.sum())
    ranges = torch.randint(0, 4, (total_draws,), generator=generator, device=device)
    lows = torch.tensor([r[0] for r in UTF8_RANGES], dtype=torch.int64, device=device)
    spans = torch.tensor([r[1] - r[0] + 1 for r in UTF8_RANGES], dtype=torch.int64, device=device)
    offsets_in_range = (torch.rand(total_draws, generator=generator, device=device, dtype=torch.float64)
                        * spans[ranges]).to(torch.int64)
    codepoints = lows[ranges] + offsets_in_range
    widths = ranges + 1
    del offsets_in_range
    owner = torch.repeat_interleave(torch.arange(len(draws), device=device), draws)
    running = torch.cumsum(widths, 0)
    starts = torch.cumsum(draws, 0) - draws
    base = torch.cat([torch.zeros(1, dtype=torch.int64, device=device), running])[starts]
    within = running - base[owner]
    keep = within <= budgets[owner]
    del running, within, base
    codepoints, widths, owner = codepoints[keep], widths[keep], owner[keep]
    string_bytes = torch.zeros(len(draws), dtype=torch.int64, device=device).index_add_(0, owner, widths)
    offsets = _torch_offsets(string_bytes)
    total = int(offsets[-1])
    out = torch.zeros(total + 1, dtype=torch.uint8, device=device)
    position = torch.cumsum(widths, 0) - widths
    for width in (1, 2, 3, 4):
        mask = widths == width
        cp, at = codepoints[mask], position[mask]
        if width == 1:
            out[at] = cp.to(torch.uint8)
            continue
        lead = {2: 0xC0, 3: 0xE0, 4: 0xF0}[width]
        out[at] = (lead | (cp >> (6 * (width - 1)))).to(torch.uint8)
        for k in range(1, width):
            out[at + k] = (0x80 | ((cp >> (6 * (width - 1 - k))) & 0x3F)).to(torch.uint8)
    return out, offsets


def mixed_c3(pairs: int = 1_000_000, long_pairs: int = 32, long_length: int = 65536, device: str = "cpu") -> Workload:
    """50% byte Levenshtein (ACGT), 25% UTF-8 Levenshtein, 25% NW (C1 matrix and gaps); lengths log-uniform
    [16, 16384] per side, plus ``long_pairs`` pairs of ``long_length`` split the same way.

    Generated with torch on ``device`` (the full batch is ~4.7 GB, so the benchmark builds it directly in HBM);
    segments hold torch tensors. The same seed gives the same strings on a given device type.
    """
    import torch

    generator = torch.Generator(device=device)
    generator.manual_seed(SEED_BASE + 3)
    b2c, costs = protein_matrix(np.random.default_rng(SEED_BASE + 1))
    kinds = torch.multinomial(torch.tensor([0.5, 0.25, 0.25], device=device), pairs, replacement=True,
                              generator=generator) if pairs else torch.zeros(0, dtype=torch.int64, device=device)
    counts = [int((kinds == k).sum()) for k in range(3)]
    long_split = [long_pairs // 2, long_pairs // 4, long_pairs - long_pairs // 2 - long_pairs // 4]

    def log_uniform(count):
        u = torch.rand(count, generator=generator, device=device, dtype=torch.float64)
        values = torch.exp(np.log(16) + u * (np.log(16385) - np.log(16))).to(torch.int64)
        return values.clamp(16, 16384)

    segments = []
    for kind_index, (name, alphabet) in enumerate([("levenshtein", DNA), ("levenshtein_utf8", None),
                                                   ("needleman_wunsch", AMINO)]):
        extra = torch.full((long_split[kind_index],), long_length, dtype=torch.int64, device=device)
        first_lengths = torch.cat([log_uniform(counts[kind_index]), extra])
        second_lengths = torch.cat([log_uniform(counts[kind_index]), extra])
        if alphabet is None:
            a, b = _torch_utf8_tape(generator, first_lengths), _torch_utf8_tape(generator, second_lengths)
        else:
            a = _torch_random_tape(generator, alphabet, first_lengths)
            b = _torch_random_tape(generator, alphabet, second_lengths)
        if name == "needleman_wunsch":
            segments.append(Segment(name, *a, *b, open=-11, extend=-1, byte_to_class=b2c, class_costs=costs))
        else:
            segments.append(Segment(name, *a, *b))
    suffix = f"_plus_{long_pairs}x{long_length}" if long_pairs else ""
    return Workload(f"mixed_{pairs}_pairs_log_uniform_16_16384{suffix}", segments)


CONFIGS = {"c0": levenshtein_c0, "c1": needleman_wunsch_c1, "c2": smith_waterman_c2, "c3": mixed_c3}
