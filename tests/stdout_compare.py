"""The reference test suite's comparison of outputs (contrib/compare/compare.pl,
used on fm_setup.out by test_suite-lsq/run_test_suite.sh:175-178), restated:
lines are lower-cased; a line containing an ignored word or matching an
ignored phrase (regular expressions, matched against the lower-cased line)
is skipped in whichever file holds it; the remaining lines are compared token
by token (whitespace-split, commas removed), numbers with a relative
tolerance of 1e-4 (|a - b| / (|a| + 1e-8)), other tokens exactly.

compare.pl stops silently at the end of the shorter file; this restatement
also requires both files to have the same number of compared lines."""
import re

IGNORE_WORDS = ["sdelta", "time", "date", "revision", "cheetah", "copyright", "parallel", "chunk", "process",
                "thread", "openmp", "memory", "mem_track"]
IGNORE_PHRASES = ["allocation limits", "clock cycles", "global chebyshev", "low gas volume", "reactant library title",
                  "replicate suffix", "startup input", "static allocations", "temperature is over the limit",
                  "the cache file", "too many iterations", "version tag", "Date", "Warning: r < rmin",
                  "Penalty potential", "..options read.", "..xmat read.", "..yvec read.", "..xmat normalized.",
                  "..yvec normalized.", "..reading single", "..weights read", "cholesky error test",
                  "..reading split", "finished", "using the lasso algorithm", "srun -N", "srun -n", "ibrun", "TACC",
                  "tacc"]
NUMBER = re.compile(r"^[-+]?[0-9]*\.?[0-9]+([eE][-+]?[0-9]+)?$")
EPS = 1.0e-4


def _kept(text):
    out = []
    for n, line in enumerate(text.split("\n")):
        low = line.lower()
        if any(re.search(w, low) for w in IGNORE_WORDS) or any(re.search(p, low) for p in IGNORE_PHRASES):
            continue
        out.append((n + 1, low))
    # the last newline ends the last line rather than starting an empty one
    if out and out[-1][1] == "" and text.endswith("\n"):
        out.pop()
    return out


def _same(t1, t2):
    if len(t1) != len(t2):
        return False
    for a, b in zip(t1, t2):
        a, b = a.replace(",", ""), b.replace(",", "")
        if a == b:
            continue
        if NUMBER.match(a) and NUMBER.match(b):
            x, y = float(a), float(b)
            if abs(x - y) / (abs(x) + 1.0e-8) < EPS:
                continue
        return False
    return True


def differences(reference, ours, limit=5):
    """Up to `limit` differing line pairs ((ref line no, text), (our line no, text));
    empty when the outputs agree."""
    r, o = _kept(reference), _kept(ours)
    bad = []
    for (nr, lr), (no, lo) in zip(r, o):
        if not _same(lr.split(), lo.split()):
            bad.append(((nr, lr), (no, lo)))
            if len(bad) >= limit:
                return bad
    if len(r) != len(o):
        k = min(len(r), len(o))
        bad.append((r[k] if k < len(r) else (None, "<end>"), o[k] if k < len(o) else (None, "<end>")))
    return bad


# Runs whose kept fm_setup.out comes from an older build than the reference
# sources: it printed the minimum pair distances with 3 decimals instead of 8
# (the same runs' params.header also lacks the EXCLD lines of the sources).
STALE_MIN_DISTANCE_DECIMALS = {"h2o-2bcheby2": 3, "skip_frames": 3}
_STAT_LINE = re.compile(r"^\t\t\d+\t")
_DECIMAL = re.compile(r"(?<![\w.])(\d+\.\d+)(?![\w.])")


def round_min_distances(text, decimals):
    """`text` with the minimum-distance values of the closing statistics
    (pair, triplet and quadruplet lines) rounded to `decimals`."""
    out = []
    for line in text.split("\n"):
        if _STAT_LINE.match(line):
            line = _DECIMAL.sub(lambda m: "%.*f" % (decimals, float(m.group(1))), line)
        out.append(line)
    return "\n".join(out)
