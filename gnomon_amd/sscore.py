"""The .sscore file of ``gnomon score`` (score/main.rs:1386-1548).

Layout, as the reference writes it:

* when regions are given, a metadata block ``#REGION\\tSCORE\\tINTERVAL``
  followed by ``#REGION\\t{name}\\t{interval}`` for every score that has one,
  in score order (main.rs:1455-1466);
* the header ``#IID`` then ``\\t{name}_AVG\\t{name}_MISSING_PCT`` per score
  (main.rs:1469-1473);
* one line per person, in output order: the IID, then per score the average
  (f64) and the missing percentage (f32) (main.rs:1499-1531), with
  AVG = sum / (count_k - missing) (0 when no variant is used) and
  MISSING_PCT = missing / count_k * 100 in f32 (0 when the score has no
  variants), the same column math as ``prep.sscore_columns``.

Numbers are written the way the ``ryu`` crate's ``Buffer::format`` prints
them (ryu 1.0.23 per the reference's Cargo.lock; the crate is not part of
the reference tree).  Its published algorithm: the shortest decimal digits
that round-trip through the type (f64 for AVG, f32 for MISSING_PCT), laid
out from the digit string d (length n, no trailing zeros) and exponent k
(value = d * 10^k), with kk = n + k, as

* ``d`` + zeros + ``.0``       when 0 <= k and kk <= L   (1234e7  -> 12340000000.0)
* ``d[:kk] . d[kk:]``           when 0 < kk <= L          (1234e-2 -> 12.34)
* ``0.`` + zeros + ``d``        when E < kk <= 0          (1234e-6 -> 0.001234)
* ``d`` ``e`` (kk-1)            when n == 1                (1e30)
* ``d[0] . d[1:]`` ``e`` (kk-1) otherwise                 (1234e30 -> 1.234e33)

with L = 16, E = -5 for f64 and L = 13, E = -6 for f32; zero is ``0.0``
(``-0.0`` with the sign bit), non-finite values ``NaN``, ``inf``, ``-inf``,
and exponents carry no ``+`` and no leading zeros.  The shortest digits come
from Python's float repr (f64) and numpy's unique float32 formatting, both
shortest-round-trip algorithms like ryu's.

The writer never leaves a partial file at the target path: the lines go to a
fresh temporary file that tempfile creates next to the target, which is
synced and then atomically renamed over it (the reference likewise writes
through a temporary file and a rename, main.rs:1409-1436, 1541-1547; the
naming and error handling here are this module's own).
"""
from __future__ import annotations

import math
import os
import tempfile
from typing import Iterable, Mapping, Sequence

import numpy as np

from .prep import sscore_columns

_LIMITS = {64: (16, -5), 32: (13, -6)}


def _digits(x: float, bits: int) -> tuple[str, int]:
    """Shortest round-trip digits (no trailing zeros) and exponent of |x| > 0."""
    if bits == 64:
        s = repr(float(x))
    else:
        s = np.format_float_scientific(np.float32(x), unique=True, trim="-")
    s = s.lstrip("-")
    mant, _, exp = s.partition("e")
    exp = int(exp) if exp else 0
    ip, _, fp = mant.partition(".")
    digits = ip + fp
    k = exp - len(fp)
    stripped = digits.lstrip("0")
    digits = stripped.rstrip("0")
    k += len(stripped) - len(digits)
    return digits, k


def format_float(x: float, bits: int = 64) -> str:
    """``ryu::Buffer::format`` of an f64 (bits=64) or f32 (bits=32) value."""
    if bits == 32:
        x = float(np.float32(x))
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    if x == 0.0:
        return sign + "0.0"
    digits, k = _digits(x, bits)
    n = len(digits)
    kk = n + k
    lim, low = _LIMITS[bits]
    if 0 <= k and kk <= lim:
        body = digits + "0" * (kk - n) + ".0"
    elif 0 < kk <= lim:
        body = digits[:kk] + "." + digits[kk:]
    elif low < kk <= 0:
        body = "0." + "0" * (-kk) + digits
    elif n == 1:
        body = digits + "e" + str(kk - 1)
    else:
        body = digits[0] + "." + digits[1:] + "e" + str(kk - 1)
    return sign + body


def sscore_lines(iids: Sequence[str], score_names: Sequence[str], score_variant_counts, sums, missing_counts,
                 regions: Mapping[str, str] | None = None) -> Iterable[str]:
    """The lines of the .sscore file (without newlines)."""
    sums = np.asarray(sums, dtype=np.float64)
    missing_counts = np.asarray(missing_counts, dtype=np.uint32)
    K = len(score_names)
    if sums.shape != (len(iids), K) or missing_counts.shape != (len(iids), K):
        raise ValueError(f"sums and missing_counts must both be {len(iids)} x {K} (people x scores); got "
                         f"{sums.shape} and {missing_counts.shape}")
    if regions:
        header_done = False
        for name in score_names:
            if name in regions:
                if not header_done:
                    yield "#REGION\tSCORE\tINTERVAL"
                    header_done = True
                yield f"#REGION\t{name}\t{regions[name]}"
    yield "#IID" + "".join(f"\t{n}_AVG\t{n}_MISSING_PCT" for n in score_names)
    avg, pct = sscore_columns(sums, missing_counts, np.asarray(score_variant_counts))
    for p, iid in enumerate(iids):
        yield str(iid) + "".join(
            f"\t{format_float(avg[p, k], 64)}\t{format_float(pct[p, k], 32)}" for k in range(K))


def write_sscore(path, iids, score_names, score_variant_counts, sums, missing_counts, regions=None) -> None:
    """Write the .sscore file at ``path``.  The text is formatted completely
    before the target directory is touched, written to a temporary file in
    that directory, synced, and renamed onto ``path``; on any failure the
    temporary file is removed and ``path`` is left as it was."""
    path = os.fspath(path)
    target_dir, base = os.path.split(os.path.abspath(path))
    if not os.path.basename(path):
        raise IsADirectoryError(f"{path!r} names a directory, not an output file")
    text = "".join(line + "\n" for line in sscore_lines(iids, score_names, score_variant_counts, sums,
                                                        missing_counts, regions))
    fd, tmp = tempfile.mkstemp(prefix=f".{base}.", suffix=".partial", dir=target_dir)
    try:
        with os.fdopen(fd, "w", encoding="utf-8", newline="\n") as f:
            f.write(text)
            f.flush()
            os.fsync(f.fileno())
        os.chmod(tmp, 0o644)
        os.replace(tmp, path)
    except BaseException:
        try:
            os.unlink(tmp)
        except FileNotFoundError:
            pass
        raise
