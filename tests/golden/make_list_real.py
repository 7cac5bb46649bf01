"""Fixture of flang's list-directed output of REAL(8) values (the editing
the reference's Tecplot writer produces, spray-simulation/cfd_outp.f90),
written by a flang-built probe of this project on values chosen to cover
every branch: random bit patterns, short decimals, powers of ten and the
edges of the F/E choice (0.1, 1e15 and their one-digit roundings), signs.

    python tests/golden/make_list_real.py   # needs flang (development container)

Output: tests/golden/flang_list_real.npz with `bits` (uint64) and `text`
(the line flang printed for each, without the leading blank)."""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

PROBE = """
program probe
  integer(8)::n,i
  real(8),allocatable::v(:)
  open(10,file='vals.bin',access='stream',form='unformatted',status='old')
  read(10)n
  allocate(v(n))
  read(10)v
  close(10)
  do i=1,n
    write(*,*)v(i)
  end do
end program
"""


def values():
    rng = np.random.default_rng(0x5EED00B1)
    bits = rng.integers(0, 2 ** 63 - 1, 6000, dtype=np.int64).astype(np.uint64)
    bits |= (rng.integers(0, 2, 6000).astype(np.uint64) << np.uint64(63))
    v = bits.view(np.float64)
    v = v[np.isfinite(v)]
    short = np.round(rng.standard_normal(4000) * 10.0 ** rng.integers(-20, 20, 4000), 3)
    mags = rng.uniform(-1, 1, 4000) * 10.0 ** rng.uniform(-6, 18, 4000)
    edges = []
    for e in range(-25, 26):
        p = 10.0 ** e
        edges += [p, np.nextafter(p, 0), np.nextafter(p, np.inf), 0.95 * p, 0.949 * p, 0.9499999 * p,
                  0.95000001 * p, 0.5 * p, 0.05 * p, 9.5 * p, 9.49 * p]
    special = [0.0, -0.0, 1.0, -1.0, 0.1, 1.0 / 3.0, 2.0 ** -1074, 2.0 ** -1022, 1.7976931348623157e308,
               1e15, 9.5e14, 9.4e14, 0.095, 0.0949, 123456789012345.6, 1e-300]
    allv = np.concatenate([v, short, mags, np.array(edges), np.array(special)])
    return allv[np.isfinite(allv)]


def main():
    v = values()
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "probe.f90"), "w") as f:
            f.write(PROBE)
        subprocess.run(["flang", "-O2", "probe.f90", "-o", "probe"], cwd=d, check=True)
        with open(os.path.join(d, "vals.bin"), "wb") as f:
            np.array([len(v)], np.int64).tofile(f)
            v.astype(np.float64).tofile(f)
        out = subprocess.run(["./probe"], cwd=d, check=True, capture_output=True, text=True).stdout
    lines = out.splitlines()
    assert len(lines) == len(v)
    assert all(l.startswith(" ") for l in lines)
    np.savez_compressed(os.path.join(HERE, "flang_list_real.npz"), bits=v.view(np.uint64),
                        text=np.array([l[1:] for l in lines]))
    print("wrote", len(v), "values")


if __name__ == "__main__":
    main()
