"""Graded mesh-line generation: the CSXCAD-style smoother that the
reference's script API uses to build rectilinear meshes
(app/ems-api.mjs:11-201: Unique, CheckSymmetry, SmoothRange,
SmoothMeshLines).

The line positions decide the grid (and with it every coefficient), so the
steps follow the reference's algorithm, including its quirks:
- near-duplicate removal drops the first line of every pair closer than
  1e-7 of the mean spacing;
- a symmetric line set (about its centre) is smoothed on one half and
  mirrored; an even set also fills the gap across the centre;
- the widest-first refinement repeatedly smooths the narrowest gap that is
  still wider than max_res, seeding its end resolutions from the
  neighbouring gaps;
- in the two-sided taper without room for a max_res middle, the stop side
  grows by the start side's resolution when start_res > stop_res (the
  reference's SmoothRange does so; kept so that meshes match line for line).
"""
import math

import numpy as np


def unique_lines(lines, tol=1e-7):
    """Sorted distinct lines with near-duplicates (spacing below tol x the
    mean spacing) reduced to the later line of the pair."""
    l = sorted(set(float(v) for v in lines))
    if len(l) < 2:
        return l
    d = [l[i + 1] - l[i] for i in range(len(l) - 1)]
    mean = sum(d) / len(d)
    drop = {i for i, v in enumerate(d) if v < mean * tol}
    return [v for i, v in enumerate(l) if i not in drop]


def check_symmetry(lines, tol=1e-10):
    """0: not symmetric about the centre; 1: symmetric with an odd count
    (the centre is a line); 2: symmetric with an even count."""
    n = len(lines)
    if n <= 2:
        return 0
    span = lines[-1] - lines[0]
    c = 0.5 * (lines[-1] + lines[0])
    for k in range(n // 2):
        if abs((c - lines[k]) - (lines[n - k - 1] - c)) > span * tol:
            return 0
    if n % 2 == 1 and abs(lines[n // 2] - c) > span * tol:
        return 0
    return 2 if n % 2 == 0 else 1


def _taper_one_side(rng, res0, ratio, max_res):
    """Offsets 0..rng growing from res0 by `ratio` up to max_res, then
    max_res steps, rescaled to end exactly at rng."""
    res, pos, n = res0, 0.0, 0
    while res < max_res and pos < rng:
        res *= ratio
        pos += res
        n += 1
    if pos > rng:
        out = [0.0]
        for m in range(1, n + 1):
            out.append(sum(res0 * ratio ** k for k in range(1, m + 1)))
        return [v * rng / pos for v in out]
    r = math.exp((math.log(max_res) - math.log(res0)) / n)
    out, pos, res = [0.0], 0.0, res0
    for _ in range(n):
        res *= r
        pos += res
        out.append(pos)
    while pos < rng:
        pos += max_res
        out.append(pos)
    return [v * rng / out[-1] for v in out]


def _grow(res0, ratio, max_res):
    """Steps needed to grow res0 to max_res by `ratio`, the exact ratio
    that reaches it in that many steps, and the distance covered."""
    res, n = res0, 0
    while res < max_res:
        res *= ratio
        n += 1
    r = math.exp((math.log(max_res) - math.log(res0)) / n)
    return n, r, sum(res0 * r ** k for k in range(1, n + 1))


def _join(start, stop, offsets, length):
    c = unique_lines(offsets)
    rng = stop - start
    tmp = [start + v * rng / length for v in c]
    return [start] + tmp[1:-1] + [stop]


def smooth_range(start, stop, start_res, stop_res, max_res, ratio):
    """Lines from start to stop, no step above max_res, growing from the
    resolution at each end by at most `ratio` per step."""
    assert ratio > 1
    rng = stop - start
    if rng < max_res and rng < start_res * ratio and rng < stop_res * ratio:
        return unique_lines([start, stop])
    lim = max_res / ratio
    if start_res >= lim and stop_res >= lim:
        n = math.ceil(rng / max_res)
        tmp = [start + (stop - start) * i / n for i in range(n + 1)]
        return [start] + tmp[1:-1] + [stop]
    if start_res < lim <= stop_res:
        tmp = [start + v for v in _taper_one_side(rng, start_res, ratio, max_res)]
        return [start] + tmp[1:-1] + [stop]
    if stop_res < lim <= start_res:
        tmp = sorted(stop - v for v in _taper_one_side(rng, stop_res, ratio, max_res))
        return [start] + tmp[1:-1] + [stop]
    n1, r1, pos1 = _grow(start_res, ratio, max_res)
    n2, r2, pos2 = _grow(stop_res, ratio, max_res)
    if pos1 + pos2 < rng:
        left = [0.0]
        for k in range(1, n1 + 1):
            left.append(left[-1] + start_res * r1 ** k)
        right = [0.0]
        for k in range(1, n2 + 1):
            right.append(right[-1] + stop_res * r2 ** k)
        for _ in range(math.ceil((rng - pos1 - pos2) / max_res)):
            left.append(left[-1] + max_res)
        length = left[-1] + right[-1]
        return _join(start, stop, left + [length - v for v in right], length)
    left, right = [0.0], [0.0]
    while left[-1] + right[-1] < rng:
        if start_res == stop_res:
            start_res *= ratio
            left.append(left[-1] + start_res)
            stop_res *= ratio
            right.append(right[-1] + start_res)
        elif start_res < stop_res:
            start_res *= ratio
            left.append(left[-1] + start_res)
        else:
            stop_res *= ratio
            right.append(right[-1] + start_res)
    length = left[-1] + right[-1]
    return _join(start, stop, left + [length - v for v in right], length)


def smooth_mesh_lines(lines, max_res, ratio=1.5):
    """Refine a set of fixed lines until no spacing exceeds max_res, with
    neighbouring spacings growing by at most `ratio`."""
    out = unique_lines(lines)
    sym = check_symmetry(out)
    centre = 0.5 * (out[-1] + out[0]) if sym else None
    if sym == 1:
        out = out[:len(out) // 2 + 1]
    elif sym == 2:
        out = out[:len(out) // 2]
    dl = [out[i + 1] - out[i] for i in range(len(out) - 1)]
    while any(v > max_res for v in dl):
        n_before = len(out)
        big = 2 * max(dl)
        mod = [big if v <= max_res else v for v in dl]
        idx = min(range(len(mod)), key=lambda i: (mod[i], i))
        sres = dl[idx - 1] if idx > 0 else max_res
        eres = dl[idx + 1] if idx < len(dl) - 1 else max_res
        out = unique_lines(out + smooth_range(out[idx], out[idx + 1], sres, eres, max_res, ratio))
        dl = [out[i + 1] - out[i] for i in range(len(out) - 1)]
        if len(out) == n_before:
            break
    if sym == 1:
        return unique_lines(out + [2 * centre - v for v in out[:-1]])
    if sym == 2:
        mid = smooth_range(out[-1], 2 * centre - out[-1], dl[-1], dl[-1], max_res, ratio)
        return unique_lines(out + mid + [2 * centre - v for v in out])
    return unique_lines(out)


def grid_lines(fixed, max_res, ratio, extra=()):
    """The script API's pattern: smooth the fixed lines, then add lines that
    must stay exact (feeds, ports) unsmoothed.  The grid is the sorted set
    with near-duplicates merged as unique_lines does (a mirrored centre line
    can land a rounding error away from an added exact line)."""
    return np.array(unique_lines(list(smooth_mesh_lines(fixed, max_res, ratio)) + [float(v) for v in extra]))
