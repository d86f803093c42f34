"""Writes tests/golden/fortran_fmt.json: how the reference's compiler (ROCm
flang) prints the edit descriptors of the reference's ASCII output files,
for edge values and a seeded sample of doubles, so that adcirc_amd.outputs
can be checked without a Fortran compiler:

- ``(2X,1PE20.10E3)``   node values (src/globalio.F:802-946),
- ``(1X,E15.7E3)``      DTDP*NSPOOL in header line 2 (src/globalio.F:1011-1051),
- ``(1X,I10,1X,I10,1X,E15.7E3,1X,I8,1X,I5,1X,'FileFmtVersion: ',I10)``,
- ``(2X,1PE20.10E3,5X,I10)``  record header (src/globalio.F:1101).

    python tests/golden/make_fmt_golden.py     (needs /opt/rocm/lib/llvm/bin/flang)
"""
import json
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FLANG = "/opt/rocm/lib/llvm/bin/flang"

SRC = """
program fmt
  implicit none
  real(8) :: v
  integer(8) :: b
  integer :: n, i, k, i1, i2, i3, i4, i5
  read(*, *) n
  do i = 1, n
    read(*, *) b
    v = transfer(b, 1.0d0)
    write(*, '(2X,1PE20.10E3)') v
    write(*, '(1X,E15.7E3)') v
  end do
  read(*, *) n
  do i = 1, n
    read(*, *) i1, i2, b, i3, i4, i5
    v = transfer(b, 1.0d0)
    write(*, 1010) i1, i2, v, i3, i4, i5
  end do
  read(*, *) n
  do i = 1, n
    read(*, *) b, k
    v = transfer(b, 1.0d0)
    write(*, '(2X,1PE20.10E3,5X,I10)') v, k
  end do
1010 FORMAT(1X,I10,1X,I10,1X,E15.7E3,1X,I8,1X,I5,1X,'FileFmtVersion: ',I10)
end program
"""


def bits(v: float) -> int:
    return struct.unpack("<q", struct.pack("<d", v))[0]


def main() -> int:
    rng = np.random.default_rng(2024)
    vals = [0.0, -0.0, 1.0, -1.0, 1234.5, 9999.99999999995, 1e-300, 1e300, 5e-324, -99999.0,
            0.1, 2.0 / 3.0, -1.2345678901234e-5, 600.0, 0.5, 1e-3, 123456789012.0,
            0.99999999999, 9.99999995, 1.00000000005, 2.5e-11, -3.75]
    mags = 10.0 ** rng.uniform(-12, 12, 400)
    vals += list(mags * rng.choice([-1.0, 1.0], 400))
    vals += list(rng.normal(0.0, 1.0, 200))
    heads = [(1, 5151, 600.0, 300, 1, 1050624), (144, 999151, 3600.0, 3600, 2, 1050624),
             (2, 861, 1.0, 1, 1, 1050624), (0, 12, 0.0, 0, 1, 1050624)]
    recs = [(600.0, 300), (86400.0, 43200), (3600.0, 3600), (1.5, 1), (21600.0, 21600)]
    inp = [str(len(vals))] + [str(bits(v)) for v in vals]
    inp += [str(len(heads))] + [f"{a} {b} {bits(c)} {d} {e} {f}" for a, b, c, d, e, f in heads]
    inp += [str(len(recs))] + [f"{bits(t)} {k}" for t, k in recs]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "fmt.f90"), os.path.join(d, "fmt")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.run([FLANG, "-O2", src, "-o", exe], check=True)
        out = subprocess.run([exe], input="\n".join(inp) + "\n", capture_output=True, text=True,
                             check=True).stdout.split("\n")
    p = 0
    e20, e15 = [], []
    for _ in vals:
        e20.append(out[p]); e15.append(out[p + 1]); p += 2
    hl = out[p:p + len(heads)]; p += len(heads)
    rl = out[p:p + len(recs)]
    res = {"bits": [bits(v) for v in vals], "e20": e20, "e15": e15,
           "headers": [list(h) for h in heads], "header_lines": hl,
           "records": [list(r) for r in recs], "record_lines": rl}
    with open(os.path.join(HERE, "fortran_fmt.json"), "w") as f:
        json.dump(res, f, indent=0)
    print(f"wrote {len(vals)} values, {len(heads)} headers, {len(recs)} record headers")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
