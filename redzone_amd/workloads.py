"""Seeded synthetic inputs for the Redzone workloads.

Everything is integer-exact torch code, so the same call gives the same
tensor on the CPU and on the GPU (except the float payload field of the
16-byte records and the tail-heavy keys, which use log/cos and are
deterministic per device).  Inputs for a parity check are generated once and
handed to both the device path and the oracle.

Removal patterns (k distinct indices in [0, n), returned as int32):
  uniform     k-prefix of a keyed pseudo-random permutation of [0, n)
              (cycle-walking bijection), i.e. uniform without replacement,
              in random order (the reference draws the same kind of list by a
              partial Fisher-Yates, DataStruct.hxx:86-96).
  clustered   [0, n) is cut into runs whose starts are chosen independently
              with probability 1/mean (geometric lengths, mean `mean`); runs
              are taken in pseudo-random order until k positions are covered;
              each run is listed ascending and the last one is truncated.
  tail_heavy  positions distributed like floor(n * Beta(5, 1)) drawn without
              replacement (draw-and-discard-repeats), produced exactly by
              weighted sampling without replacement with the discrete
              Beta(5,1) masses w_p ~ (p+1)^5 - p^5 (exponential race).
  redzone_pct the reference's red-zone-share lists (redzone_pct_removals).
"""
import math

import torch

MASK32 = 0xFFFFFFFF
_ROUND_MULS = (0x2545F491, 0x4C957F2D, 0x1B873593, 0x68E31DA5)  # odd, < 2^31
_HASH_MULS = (0x045D9F3B, 0x0293A5D1)  # odd, < 2^27 (products stay below 2^59)


def _as_i64(x):
    return x.to(torch.int64)


def hash32(x, key):
    """32-bit integer hash of int64 values in [0, 2^32); exact in int64."""
    x = (x ^ (key & MASK32)) & MASK32
    x = ((x ^ (x >> 16)) * _HASH_MULS[0]) & MASK32
    x = ((x ^ (x >> 15)) * _HASH_MULS[1]) & MASK32
    return x ^ (x >> 16)


def _round_keys(seed, salt):
    keys = []
    s = (seed * 0x9E3779B1 + salt) & MASK32
    for r in range(4):
        s = (s * 1664525 + 1013904223) & MASK32
        keys.append(s)
    return keys


def permute_bits(x, bits, seed, salt=0):
    """Keyed bijection on [0, 2^bits), bits <= 32 (add, odd multiply, xorshift rounds)."""
    mask = (1 << bits) - 1
    sh = max(1, bits // 2)
    for r, key in enumerate(_round_keys(seed, salt)):
        x = (x + (key & mask)) & mask
        x = (x * _ROUND_MULS[r]) & mask
        x = x ^ (x >> sh)
    return x


def permute_range(i, n, seed, salt=0):
    """Keyed bijection on [0, n) by cycle walking over the next power of two."""
    bits = max(1, int(n - 1).bit_length())
    y = permute_bits(i, bits, seed, salt)
    while True:
        over = y >= n
        if not bool(over.any()):
            return y
        y = torch.where(over, permute_bits(y, bits, seed, salt), y)


def _to_int32(x):
    """uint32 values held in int64 -> int32 with the same bits."""
    return torch.where(x >= 2 ** 31, x - 2 ** 32, x).to(torch.int32)


# --------------------------------------------------------------------------
# removal lists
# --------------------------------------------------------------------------

def uniform_removals(n, k, seed, device="cpu"):
    if k == 0:
        return torch.empty(0, dtype=torch.int32, device=device)
    i = torch.arange(k, dtype=torch.int64, device=device)
    return permute_range(i, n, seed, salt=1).to(torch.int32)


def clustered_removals(n, k, seed, mean=256, device="cpu"):
    if k == 0:
        return torch.empty(0, dtype=torch.int32, device=device)
    p = torch.arange(n, dtype=torch.int64, device=device)
    start = (hash32(p, seed * 7 + 3) % mean) == 0
    start[0] = True
    starts = torch.nonzero(start).flatten()
    del p, start
    m = starts.numel()
    lengths = torch.diff(starts, append=torch.tensor([n], dtype=torch.int64, device=device))
    order = permute_range(torch.arange(m, dtype=torch.int64, device=device), m, seed, salt=2)
    lo = lengths[order]
    cum = torch.cumsum(lo, 0)
    t = int(torch.searchsorted(cum, torch.tensor([k], dtype=torch.int64, device=device)).item())
    sel = order[: t + 1]
    lsel = lo[: t + 1]
    excl = cum[: t + 1] - lsel
    total = int(cum[t].item())
    idx = torch.arange(total, dtype=torch.int64, device=device)
    R = torch.repeat_interleave(starts[sel], lsel) + idx - torch.repeat_interleave(excl, lsel)
    return R[:k].to(torch.int32)


def tail_heavy_removals(n, k, seed, device="cpu"):
    if k == 0:
        return torch.empty(0, dtype=torch.int32, device=device)
    p = torch.arange(n, dtype=torch.int64, device=device)
    u = (hash32(p, seed * 11 + 5).to(torch.float64) + 0.5) / 4294967296.0
    pf = p.to(torch.float64)
    logw = torch.log((((5.0 * pf + 10.0) * pf + 10.0) * pf + 5.0) * pf + 1.0)  # (p+1)^5 - p^5 > 0
    key = torch.log(-torch.log(u)) - logw  # log of the arrival time Exp(1)/w_p
    del u, pf, logw
    if k * 16 <= n:
        R = torch.topk(key, k, largest=False, sorted=True).indices
    else:
        R = torch.sort(key, stable=True).indices[:k]
    return R.to(torch.int32)


def redzone_pct_removals(n, k, red_pct, seed, device="cpu"):
    """The reference's test lists with a red-zone share (TestLists_t::Update(n,
    k, r), DataStruct.hxx:97-122): red_k = floor(k * r / 100) removals drawn
    uniformly without replacement from the red zone [n-k, n) and the other
    k - red_k from the green zone [0, n-k), listed green first, then red.
    r = -1 is the plain uniform list over [0, n) (DataStruct.hxx:85-95).  The
    reference sweeps r over {100, -1, 0, 25, 50, 75} (redzone.hxx:1094) and
    skips r >= 0 when k > n/2 (redzone.hxx:1121), where the green zone is
    smaller than k - red_k."""
    if red_pct == -1:
        return uniform_removals(n, k, seed, device)
    if not 0 <= red_pct <= 100:
        raise ValueError("red_pct is -1 or a percentage")
    red = (k * red_pct) // 100
    green = k - red
    if green > n - k:
        raise ValueError("the green zone [0, n-k) holds fewer than k - red_k positions")
    parts = []
    if green:
        parts.append(permute_range(torch.arange(green, dtype=torch.int64, device=device), n - k, seed, salt=7))
    if red:
        parts.append((n - k) + permute_range(torch.arange(red, dtype=torch.int64, device=device), k, seed, salt=8))
    if not parts:
        return torch.empty(0, dtype=torch.int32, device=device)
    return torch.cat(parts).to(torch.int32)


PATTERNS = {"uniform": uniform_removals, "clustered": clustered_removals, "tail_heavy": tail_heavy_removals}


def removals(pattern, n, k, seed, device="cpu", **kw):
    return PATTERNS[pattern](n, k, seed, device=device, **kw)


# --------------------------------------------------------------------------
# payloads
# --------------------------------------------------------------------------

def index_payload(n, device="cpu"):
    """int32 A[i] = i (the reference's test list, DataStruct.hxx:77)."""
    return torch.arange(n, dtype=torch.int32, device=device)


def unique_u32_payload(n, seed, device="cpu"):
    """n unique uint32 payloads (a keyed bijection of the index on 32 bits), as int32 bits."""
    i = torch.arange(n, dtype=torch.int64, device=device)
    return _to_int32(permute_bits(i, 32, seed, salt=3))


def u64_payload(n, seed, device="cpu"):
    """n unique 8-byte items as an (n, 2) int32 tensor: low word unique, high word hashed."""
    i = torch.arange(n, dtype=torch.int64, device=device)
    lo = _to_int32(permute_bits(i, 32, seed, salt=4))
    hi = _to_int32(hash32(i, seed * 13 + 1))
    return torch.stack([lo, hi], dim=1).contiguous()


def rec16_payload(n, seed, device="cpu"):
    """n 16-byte records {u64 key unique, u32 tag uniform, f32 val ~ N(0,1)} as an (n, 4) int32 tensor."""
    i = torch.arange(n, dtype=torch.int64, device=device)
    key_lo = _to_int32(permute_bits(i, 32, seed, salt=5))
    key_hi = _to_int32(hash32(i, seed * 17 + 2))
    tag = _to_int32(hash32(i, seed * 19 + 9))
    u1 = (hash32(i, seed * 23 + 4).to(torch.float64) + 0.5) / 4294967296.0
    u2 = (hash32(i, seed * 29 + 6).to(torch.float64) + 0.5) / 4294967296.0
    val = (torch.sqrt(-2.0 * torch.log(u1)) * torch.cos(2.0 * math.pi * u2)).to(torch.float32)
    return torch.stack([key_lo, key_hi, tag, val.view(torch.int32)], dim=1).contiguous()


# 32-bit words that a marking scheme could mistake for a mark: the reference's
# GPU sentinel 0x80000000 (redzone.hxx:704, also the bits of float -0.0), its
# O(n) marker -1 (redzone.hxx:594), zero, and float NaNs with payloads, the
# infinities and a denormal
EDGE_WORDS = (0x80000000, 0xFFFFFFFF, 0x00000000, 0x7FFFFFFF, 0x7FC00001, 0x7F800001, 0xFFC12345, 0x7F800000,
              0xFF800000, 0x00000001)


def edge_payload(elem_bytes, n, seed, device="cpu"):
    """n items of elem_bytes (an (n, w) int32 tensor, or (n,) for 4 bytes) in
    which every word of about half of the items is one of EDGE_WORDS (so the
    sentinels sit in the head and in the red zone alike) and the rest are
    unique hashes; the first word of an item is unique whenever it is not an
    edge word.  Moved items must keep these bits exactly."""
    w = elem_bytes // 4
    i = torch.arange(n, dtype=torch.int64, device=device)
    edge = torch.tensor(EDGE_WORDS, dtype=torch.int64, device=device)
    cols = []
    for c in range(w):
        h = hash32(i, seed * 31 + 7 * c + 1)
        pick = edge[h % len(EDGE_WORDS)]
        plain = permute_bits(i, 32, seed, salt=9 + c) if c == 0 else hash32(i, seed * 37 + c)
        cols.append(_to_int32(torch.where((h >> 8) % 2 == 0, pick, plain)))
    out = torch.stack(cols, dim=1).contiguous()
    return out.reshape(n) if w == 1 else out


def payload(elem_bytes, n, seed, device="cpu"):
    if elem_bytes == 4:
        return unique_u32_payload(n, seed, device)
    if elem_bytes == 8:
        return u64_payload(n, seed, device)
    if elem_bytes == 16:
        return rec16_payload(n, seed, device)
    raise ValueError(f"unsupported element size {elem_bytes}")
