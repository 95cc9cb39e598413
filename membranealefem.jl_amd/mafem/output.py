"""Text files of an analysis, in the reference's formats.

With `p.output` the reference writes, under `out_path`:
  * `t<id>-xms.txt`, `t<id>-cps.txt`: `writedlm` of the state matrices, one
    tab-separated row per control point (src/Input.jl:111-114,
    src/Analysis.jl:75-78); `restart` reads them back with `readdlm`
    (src/MembraneAleFem.jl:90-95, src/Input.jl:35-37);
  * `times.txt`: "<n> times:" then "<t_id>\\t<time>" per step
    (src/Analysis.jl:37-44);
  * `f-pull.txt` (F_PULL): the header "time\\txp\\typ\\tzp\\tfx\\tfy\\tfz", then
    one line per step with the time, the pulled element's mean position and
    the pull force (src/Analysis.jl:47-56, 81-90);
  * `out_file`: progress lines, among them "--> iteration k: ε = e" per
    Newton iteration (src/analysis/FiniteElement.jl:30-51).

Numbers are written as Julia prints a Float64: the shortest decimal string
that reads back to the same double, in plain notation when the position pt of
the decimal point relative to the first significant digit satisfies
-4 < pt <= 6, else as mantissa "e" exponent (1.0e-5, 1.234567e6), with ".0"
after integral mantissas and no "+" or zero padding in the exponent.
"""
import decimal
import math
import os

import numpy as np


def julia_float(x):
    """`print(io, x::Float64)` of Julia."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    if x == 0.0:
        return sign + "0.0"
    # shortest round-trip digits (Python's repr is shortest and correctly rounded)
    t = decimal.Decimal(repr(abs(x))).normalize().as_tuple()
    digits = "".join(str(d) for d in t.digits)
    pt = t.exponent + len(digits)  # value = 0.<digits> x 10^pt
    nd = len(digits)
    if -4 < pt <= 6:
        if pt <= 0:
            body = "0." + "0" * (-pt) + digits
        elif pt >= nd:
            body = digits + "0" * (pt - nd) + ".0"
        else:
            body = digits[:pt] + "." + digits[pt:]
    else:
        body = digits[0] + "." + (digits[1:] or "0") + "e" + str(pt - 1)
    return sign + body


def writedlm(path, a):
    """`writedlm(path, A)`: rows of A, entries separated by tabs."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a[:, None]
    with open(path, "w") as io:
        for row in a:
            io.write("\t".join(julia_float(v) for v in row) + "\n")


def readdlm(path):
    """`readdlm(path)` of a numeric matrix written by writedlm."""
    return np.loadtxt(path, dtype=np.float64, ndmin=2)


def append(path, text):
    with open(path, "a") as io:
        io.write(text)


def state_files(out_path, t_id):
    return (os.path.join(out_path, f"t{t_id}-xms.txt"),
            os.path.join(out_path, f"t{t_id}-cps.txt"))


def write_state(out_path, t_id, xms, cps):
    fx, fc = state_files(out_path, t_id)
    writedlm(fx, xms)
    writedlm(fc, cps)


def times_text(times, t0_id):
    lines = [f"{len(times) + t0_id} times:"]
    lines += [f"{k + 1 + t0_id}\t{julia_float(t)}" for k, t in enumerate(times)]
    return "\n".join(lines) + "\n"


PULL_HEADER = "time\txp\typ\tzp\tfx\tfy\tfz\n"


def pull_line(time, x_pull, f_pull):
    vals = [time, *x_pull, *f_pull]
    return "\t".join(julia_float(v) for v in vals) + "\n"


def iteration_lines(eps):
    return "".join(f"--> iteration {k + 1}: ε = {julia_float(e)}\n" for k, e in enumerate(eps))
