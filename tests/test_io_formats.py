"""The reference's text outputs (hor.io): numbers as C++ setprecision(16)
writes them (C's %.16g, checked against the C library's snprintf), the
velocity dump layout (highreg_stokes_2surf/main.cpp:258-270) and the two
mid-plane slice layouts (extend_to_grid/main.cpp:294-303, 543-549)."""
import ctypes as C
import ctypes.util
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from hor import grid as HG
from hor import io as HIO

REF_MAIN = os.path.join(ROOT, "oracle", "_ref", "ref2surf_main")


def _c_fmt(v):
    libc = C.CDLL(ctypes.util.find_library("c"))
    buf = C.create_string_buffer(64)
    libc.snprintf(buf, 64, b"%.16g", C.c_double(v))
    return buf.value.decode()


def test_number_format_is_c_precision_16():
    rng = np.random.default_rng(1)
    vals = np.concatenate([rng.standard_normal(200) * 10.0 ** rng.integers(-20, 20, 200),
                           [0.0, -0.0, 1.0, -1.5, 1e-300, 123456789012345678.0, 0.1, 1 / 3]])
    for v in vals:
        assert HIO.fmt(v) == _c_fmt(v), v


def test_velocity_dump_round_trip(tmp_path):
    u = np.random.default_rng(2).standard_normal(3 * 50)
    p = tmp_path / "vel1"
    HIO.write_velocity(p, u)
    lines = p.read_text().splitlines()
    assert len(lines) == 50 and all(len(line.split(" ")) == 3 for line in lines)
    back = HIO.read_velocity(p)
    assert np.abs(back - u).max() <= 1e-15 * np.abs(u).max()


def test_velocity_dump_bytes_match_the_reference_program(tmp_path):
    """The reference two-surface program (its own sources, built under
    oracle/_ref) writes its velocities with setprecision(16)
    (main.cpp:258-270).  read_velocity then write_velocity on those dumps
    gives back the same bytes, except for numbers whose 16-digit decimal does
    not name its double (printing the parsed double gives another string):
    those come back as %.16g of that double, as the reference would print it."""
    if not os.path.exists(REF_MAIN):
        pytest.skip("reference program (oracle/_ref/ref2surf_main) not built")
    run = subprocess.run([REF_MAIN], cwd=tmp_path, capture_output=True, text=True, timeout=600)
    assert run.returncode == 0, run.stderr[-2000:]
    for name in ("vel1_N32_highreg_57_k4", "vel2_N32_highreg_57_k4"):
        ref = (tmp_path / name).read_text()
        HIO.write_velocity(tmp_path / "again", HIO.read_velocity(tmp_path / name))
        ours = (tmp_path / "again").read_text()
        lines = ref.split("\n")
        assert lines[-1] == "" and all(len(line.split(" ")) == 3 for line in lines[:-1])
        renamed = 0
        expect = []
        for line in lines[:-1]:
            toks = []
            for tok in line.split(" "):
                again = "%.16g" % float(tok)
                renamed += again != tok
                toks.append(again)
            expect.append(" ".join(toks))
        print(f"{name}: {len(lines) - 1} lines, {renamed} of {3 * (len(lines) - 1)} numbers not named by "
              f"their 16 digits", flush=True)
        assert renamed <= 0.05 * 3 * (len(lines) - 1)
        assert ours == "\n".join(expect) + "\n"
        same = sum(a == b for a, b in zip(ours.split("\n"), lines))
        assert same >= 0.9 * len(lines)


def test_mid_plane_slice_layouts(tmp_path):
    N, L = 6, 3.0
    h = L / N
    f = np.arange((N + 1) ** 3, dtype=np.float64).reshape((N + 1,) * 3) * 0.25
    pj, pi = tmp_path / "p_midY", tmp_path / "err_u_midY"
    HIO.write_mid_y_slice(pj, f, h, newline_per="j")
    HIO.write_mid_y_slice(pi, f, h, newline_per="i")
    lj, li = pj.read_text().split("\n"), pi.read_text().split("\n")
    assert len(lj) == (N + 1) ** 2 + 1 and len(li) == N + 2        # + the empty tail after the last break
    mid = N // 2                                                    # j h = 1.5
    for i in range(N + 1):
        row = " ".join(HIO.fmt(v) for v in f[i, mid]).replace(" ", "  ") + "  "
        assert li[i] == row
        assert lj[i * (N + 1) + mid] == row
        assert all(lj[i * (N + 1) + j] == "" for j in range(N + 1) if j != mid)
    assert np.array_equal(HIO.read_mid_y_slice(pi), f[:, mid, :])


def test_d4_exact_on_quartics():
    N, L = 10, 3.0
    h = L / N
    g = np.arange(N + 1) * h
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    f = X ** 4 - 2 * X * Y ** 3 + Z ** 4 * Y
    grads = (4 * X ** 3 - 2 * Y ** 3, -6 * X * Y ** 2 + Z ** 4, 4 * Z ** 3 * Y)
    for ax in range(3):
        assert np.abs(HG.d4(f, h, ax) - grads[ax]).max() <= 1e-9
