"""Particle backup files (main.f90:1102-1132 writer, restart.f90:56-113
reader) on the host: the library writes the bytes the reference's Fortran
statements write (oracle/backup_writer.f90 compiled by amdflang), reads
such files back, concatenates the partitions of a file written with several
OpenMP threads, and follows the split-record convention."""
import os
import shutil
import struct
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import jd_lepic_amd as L  # noqa: E402

FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")


def fortran_values(np_tot):
    """The arrays oracle/backup_writer.f90 fills: vxp(k,i,p,ip) = k + 8i +
    1024p + 65536ip + 0.125 (1-based)."""
    parts = []
    for ip, row in enumerate(np_tot, start=1):
        parts.append([])
        for p, n in enumerate(row, start=1):
            i = np.arange(1, n + 1)[:, None]
            k = np.arange(1, 7)[None, :]
            parts[-1].append((k + 8 * i + 1024 * p + 65536 * ip + 0.125).astype(np.float64))
    return parts


@pytest.fixture(scope="module")
def fortran_writer(tmp_path_factory):
    if not os.path.exists(FLANG):
        pytest.skip("amdflang not in this image")
    d = tmp_path_factory.mktemp("bw")
    exe = d / "backup_writer"
    subprocess.run([FLANG, "-O1", os.path.join(ROOT, "oracle", "backup_writer.f90"), "-o", str(exe)],
                   check=True, cwd=d, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    return exe


@pytest.mark.parametrize("np_tot", [[[3, 0]], [[3, 0], [1, 2]], [[0, 0, 5], [7, 1, 0], [2, 2, 2]]])
def test_bytes_match_fortran_writer(fortran_writer, tmp_path, np_tot):
    ref = tmp_path / "ref.bak"
    counts = [str(n) for row in np_tot for n in row]
    subprocess.run([str(fortran_writer), str(ref), str(len(np_tot[0])), str(len(np_tot)), *counts], check=True,
                   cwd=tmp_path)
    ours = tmp_path / "ours.bak"
    L.backup_write(ours, 1.5e-6, 7, 0.75, fortran_values(np_tot))
    assert ref.read_bytes() == ours.read_bytes()
    # and the reader takes the Fortran file
    t, g, pc, counts = L.backup_header(ref, len(np_tot[0]))
    assert (t, g, pc) == (1.5e-6, 7, 0.75)
    assert counts.tolist() == np_tot


def test_round_trip_concatenates_partitions(tmp_path):
    rng = np.random.default_rng(4)
    parts = [[rng.standard_normal((n, 6)) for n in row] for row in [[5, 0, 3], [2, 4, 0]]]
    f = tmp_path / "p.bak"
    L.backup_write(f, 3.25, 2, 1.0, parts)
    for p in range(3):
        got = L.backup_read_species(f, 3, p + 1)
        want = np.concatenate([parts[0][p], parts[1][p]])
        assert np.array_equal(got, want)


def marker(n):
    return struct.pack("<i", n)


def test_split_record_is_reassembled(tmp_path):
    """A record written as two subrecords: negative header on the first (more
    follows), negative footer on the second (one came before)."""
    v = np.arange(18, dtype=np.float64).reshape(3, 6)
    raw = v.tobytes()
    a, b = raw[:40], raw[40:]
    body = marker(20) + struct.pack("<did", 0.5, 1, 2.0) + marker(20)
    body += marker(4) + struct.pack("<i", 3) + marker(4)
    body += marker(-len(a)) + a + marker(len(a)) + marker(len(b)) + b + marker(-len(b))
    f = tmp_path / "split.bak"
    f.write_bytes(body)
    assert np.array_equal(L.backup_read_species(f, 1, 1), v)


def test_malformed_files_are_rejected(tmp_path):
    parts = [[np.ones((4, 6))]]
    good = tmp_path / "g.bak"
    L.backup_write(good, 0.0, 0, 1.0, parts)
    raw = good.read_bytes()
    cut = tmp_path / "cut.bak"
    cut.write_bytes(raw[:-10])
    with pytest.raises(L.LepicError):
        L.backup_read_species(cut, 1, 1)
    with pytest.raises(L.LepicError):          # record 2 does not hold a multiple of ntype counts
        L.backup_header(good, 2)
    bad = bytearray(raw)
    bad[-4:] = marker(12)                       # footer disagrees with header
    badf = tmp_path / "bad.bak"
    badf.write_bytes(bytes(bad))
    with pytest.raises(L.LepicError):
        L.backup_read_species(badf, 1, 1)
