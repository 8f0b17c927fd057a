"""Dump-case directories, MatrixMarket files and the cupf_bench command line.

The dump-case format is the one the reference's C++ bench reads
(cuPF/tests/dump_case_loader.cpp:92-231) and its Python tooling writes
(benchmark/common/matpower_data.py:350-358: scipy.io.mmwrite for the Ybus,
numpy.savetxt "%.18e" pairs and "%d" indices).  Directories written the way
the reference writes them (scipy / numpy here) are read by the product's
loader; the product's writer round-trips bit for bit; the reference's own
MatrixMarket Jacobians (tests/golden) read the same as scipy reads them.  On
the GPU, a dump case of the 118-bus grid solves through NewtonSolver and
through cupf_bench to the oracle's solution.
"""

from __future__ import annotations

import gzip
import subprocess
from pathlib import Path

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = Path(__file__).resolve().parent / "golden"
CLI = ROOT / "gpu-powerflow_amd" / "cupf" / "cupf_bench"


def write_like_reference(d: Path, g, sbus, v0, vref=None, symmetry=None):
    """A dump case as the reference's save_dump_case writes it."""
    d.mkdir(parents=True, exist_ok=True)
    Y = sp.csr_matrix((g.data, g.indices, g.indptr), shape=(g.n_bus, g.n_bus))
    if symmetry is None:
        scipy.io.mmwrite(str(d / "dump_Ybus.mtx"), Y)
    else:
        scipy.io.mmwrite(str(d / "dump_Ybus.mtx"), Y, symmetry=symmetry)
    np.savetxt(d / "dump_Sbus.txt", np.column_stack((sbus.real, sbus.imag)), fmt="%.18e")
    np.savetxt(d / "dump_V.txt", np.column_stack((v0.real, v0.imag)), fmt="%.18e")
    np.savetxt(d / "dump_pv.txt", g.pv.astype(np.int32), fmt="%d")
    np.savetxt(d / "dump_pq.txt", g.pq.astype(np.int32), fmt="%d")
    if vref is not None:
        np.savetxt(d / "dump_Vref.txt", np.column_stack((vref.real, vref.imag)), fmt="%.18e")


@pytest.mark.parametrize("symmetry", [None, "symmetric"])
def test_reads_a_case_written_like_the_reference(tmp_path, cupf, syn, symmetry):
    """scipy's default (general) and its symmetric form (lower triangle, the
    loader mirrors it); the grid's Ybus is symmetric."""
    g = syn.workload_grid("case118")
    S = g.sbus_base()
    V0 = g.flat_start()
    write_like_reference(tmp_path / "case118", g, S, V0, symmetry=symmetry)
    head = (tmp_path / "case118" / "dump_Ybus.mtx").read_text().splitlines()[0]
    assert "complex" in head and (symmetry or "general") in head
    c = cupf.load_dump_case(str(tmp_path / "case118"))
    assert c["case_name"] == "case118" and c["rows"] == g.n_bus and c["vref"] is None
    np.testing.assert_array_equal(c["indptr"], g.indptr)
    np.testing.assert_array_equal(c["indices"], g.indices)
    np.testing.assert_array_equal(c["data"], g.data)
    np.testing.assert_array_equal(c["sbus"], S)
    np.testing.assert_array_equal(c["v0"], V0)
    np.testing.assert_array_equal(c["pv"], g.pv)
    np.testing.assert_array_equal(c["pq"], g.pq)


def test_writer_round_trips_bit_for_bit(tmp_path, cupf, syn):
    g = syn.workload_grid("case118")
    rng = np.random.default_rng(3)
    S = g.sbus_base() * (1 + 1e-3 * rng.standard_normal(g.n_bus))
    V0 = g.flat_start() * np.exp(1j * 1e-2 * rng.standard_normal(g.n_bus))
    Vref = V0 * 0.99
    cupf.save_dump_case(str(tmp_path / "c"), g.indptr, g.indices, g.data, S, V0, g.pv, g.pq,
                        vref=Vref)
    c = cupf.load_dump_case(str(tmp_path / "c"))
    for k, v in (("indptr", g.indptr), ("indices", g.indices), ("data", g.data), ("sbus", S),
                 ("v0", V0), ("pv", g.pv), ("pq", g.pq), ("vref", Vref)):
        np.testing.assert_array_equal(c[k], v, err_msg=k)
    # and scipy reads the product's Ybus the same
    Y = scipy.io.mmread(str(tmp_path / "c" / "dump_Ybus.mtx")).tocsr()
    Y.sort_indices()
    np.testing.assert_array_equal(Y.toarray(), sp.csr_matrix(
        (g.data, g.indices, g.indptr), shape=(g.n_bus, g.n_bus)).toarray())


@pytest.mark.parametrize("case", ["case118", "case1354pegase"])
def test_reference_jacobian_files_read_as_scipy_reads_them(tmp_path, cupf, case):
    for name in ("J.mtx", "F.mtx"):
        (tmp_path / name).write_bytes(gzip.decompress((GOLDEN / case / f"{name}.gz").read_bytes()))
    J = cupf.read_matrix_market(str(tmp_path / "J.mtx"))
    Js = scipy.io.mmread(str(tmp_path / "J.mtx")).tocsr()
    Js.sum_duplicates()
    Js.sort_indices()
    assert not J["dense"] and (J["rows"], J["cols"]) == Js.shape
    np.testing.assert_array_equal(J["indptr"], Js.indptr)
    np.testing.assert_array_equal(J["indices"], Js.indices)
    np.testing.assert_array_equal(J["values"], Js.data)
    F = cupf.read_matrix_market(str(tmp_path / "F.mtx"))
    Fs = np.asarray(scipy.io.mmread(str(tmp_path / "F.mtx")))
    assert F["dense"]
    np.testing.assert_array_equal(F["values"], Fs)
    # write both back and read them again
    cupf.write_matrix_market(str(tmp_path / "J2.mtx"), J["values"], J["indptr"], J["indices"],
                             (J["rows"], J["cols"]))
    cupf.write_matrix_market(str(tmp_path / "F2.mtx"), F["values"])
    J2 = cupf.read_matrix_market(str(tmp_path / "J2.mtx"))
    np.testing.assert_array_equal(J2["values"], J["values"])
    np.testing.assert_array_equal(cupf.read_matrix_market(str(tmp_path / "F2.mtx"))["values"],
                                  F["values"])


def test_symmetric_and_duplicate_entries(tmp_path, cupf):
    (tmp_path / "s.mtx").write_text(
        "%%MatrixMarket matrix coordinate complex symmetric\n% comment\n3 3 4\n"
        "1 1 1.0 0.5\n2 1 2.0 -1.0\n3 3 3.0 0.0\n2 1 0.5 0.25\n")
    m = cupf.read_matrix_market(str(tmp_path / "s.mtx"))
    A = sp.csr_matrix((m["values"], m["indices"], m["indptr"]), shape=(3, 3)).toarray()
    expect = np.array([[1 + 0.5j, 2.5 - 0.75j, 0], [2.5 - 0.75j, 0, 0], [0, 0, 3]])
    np.testing.assert_array_equal(A, expect)


def test_malformed_inputs_raise(tmp_path, cupf):
    with pytest.raises(RuntimeError, match="does not exist"):
        cupf.load_dump_case(str(tmp_path / "missing"))
    (tmp_path / "bad.mtx").write_text("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n")
    with pytest.raises(RuntimeError, match="out of range"):
        cupf.read_matrix_market(str(tmp_path / "bad.mtx"))
    (tmp_path / "pat.mtx").write_text("%%MatrixMarket matrix coordinate pattern general\n1 1 1\n1 1\n")
    with pytest.raises(RuntimeError, match="unsupported"):
        cupf.read_matrix_market(str(tmp_path / "pat.mtx"))


def test_cli_argument_errors():
    r = subprocess.run([str(CLI)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([str(CLI), "/nonexistent", "custom-fp64", "4"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "does not exist" in r.stderr
    r = subprocess.run([str(CLI), "/nonexistent", "cudss-fp64", "4"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "not part of this build" in r.stderr


@pytest.mark.gpu
def test_dump_case_solves_on_the_gpu(gpu, tmp_path, cupf, syn, oracle):
    """A dump case of the 118-bus grid, written as the reference writes it,
    solved through NewtonSolver and through cupf_bench (solve and debug
    modes): V within 1e-9 of the oracle converged to 1e-13 (the dump's Vref)."""
    g = syn.workload_grid("case118")
    S = g.sbus_scaled(syn.load_factors(1, g.n_bus, 0.7, 1.3, seed=1))[0]
    V0 = g.flat_start()
    Vref = oracle.nr_solve_batch(g.n_bus, g.indptr, g.indices, g.data, S[None], V0[None], g.pv,
                                 g.pq, tol=1e-13)[0][0]
    d = tmp_path / "case118"
    write_like_reference(d, g, S, V0, Vref)
    c = cupf.load_dump_case(str(d))
    for compute, prec in (("FP64", "FP64"), ("Mixed", "FP32")):
        o = cupf.NewtonOptions()
        o.compute = getattr(cupf.ComputePolicy, compute)
        o.custom.precision = getattr(cupf.CustomPrecision, prec)
        s = cupf.NewtonSolver(o)
        s.initialize(c["indptr"], c["indices"], c["data"], c["rows"], c["cols"], c["pv"], c["pq"])
        r = s.solve(c["indptr"], c["indices"], c["data"], c["rows"], c["cols"], c["sbus"], c["v0"],
                    c["pv"], c["pq"])
        assert r.converged
        assert np.abs(np.angle(r.V_numpy) - np.angle(c["vref"])).max() <= 1e-9
        assert np.abs(np.abs(r.V_numpy) - np.abs(c["vref"])).max() <= 1e-9
    print("[case-io] cupf_bench solve / debug", flush=True)
    r = subprocess.run([str(CLI), str(d), "custom-fp64", "1,8", "solve", "3"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines[0].startswith("case=case118 n_bus=118")
    assert lines[1] == "B,init_ms,solve_ms,solve_per_sys_ms,converged_count,max_mismatch,max_iterations"
    for line, B in zip(lines[2:], (1, 8)):
        f = line.split(",")
        assert int(f[0]) == B and int(f[4]) == B and float(f[5]) <= 1e-6
    r = subprocess.run([str(CLI), str(d), "custom-tf32", "4", "debug"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [l.split(",") for l in r.stdout.strip().splitlines()[2:]]
    assert len(rows) == 4 and all(int(x[3]) == 1 for x in rows)
    r = subprocess.run([str(CLI), str(d), "custom-mixed", "4", "operators", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    f = r.stdout.strip().splitlines()[2].split(",")
    assert float(f[1]) > 0 and float(f[4]) > 0 and int(f[7]) == 4
