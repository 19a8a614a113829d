"""C ABI surface checks that need no GPU."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = []
    for hdr in (ROOT / "include").glob("*.h"):
        txt = re.sub(r"/\*.*?\*/", "", hdr.read_text(), flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(tx_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "tx_plan_create" in names and "tx_boxcar_run" in names and len(names) >= 18


def test_library_exports_every_declared_symbol():
    from transientx_amd import _lib

    L = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert set(_lib.EXPORTS) == set(declared_functions())
    assert L.tx_version() >= 1


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of the structs agree with the C compiler's layout."""
    import shutil
    import subprocess

    from transientx_amd import _lib

    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    structs = {"tx_beam_desc": _lib.BeamDesc, "tx_plan_desc": _lib.PlanDesc,
               "tx_info": _lib.Info, "tx_ddrow": _lib.DdRow, "tx_cluster_desc": _lib.ClusterDesc,
               "tx_cands_ctx": _lib.CandsCtx, "tx_refine_desc": _lib.RefineDesc,
               "tx_diff_desc": _lib.DiffDesc, "tx_fil_header": _lib.FilHeader}
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "txgpu.h"', "int main(void){"]
    for cname, py in structs.items():
        src.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            src.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    src.append('printf("tx_point %zu\\n", sizeof(tx_point));')
    src.append('printf("tx_cand %zu\\n", sizeof(tx_cand));')
    for f in _lib.REFINE_RESULT_DTYPE.names:
        src.append(f'printf("tx_refine_result.{f} %zu\\n", offsetof(tx_refine_result, {f}));')
    src.append('printf("tx_refine_result %zu\\n", sizeof(tx_refine_result)); return 0; }')
    (tmp_path / "l.c").write_text("\n".join(src))
    subprocess.run(["gcc", "-I", str(ROOT / "include"), "-o", str(tmp_path / "l"),
                    str(tmp_path / "l.c")], check=True)
    out = dict(line.split() for line in subprocess.run(
        [str(tmp_path / "l")], capture_output=True, text=True, check=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(out[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(out[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)
    assert int(out["tx_point"]) == _lib.POINT_DTYPE.itemsize
    assert int(out["tx_cand"]) == _lib.CAND_DTYPE.itemsize
    rr = _lib.REFINE_RESULT_DTYPE
    assert int(out["tx_refine_result"]) == rr.itemsize
    for f in rr.names:
        assert int(out[f"tx_refine_result.{f}"]) == rr.fields[f][1], f


def test_errors_come_back_as_codes_not_exceptions():
    from transientx_amd import _lib

    L = _lib.lib()
    assert L.tx_beam_create(None, None) == _lib.TX_EINVAL
    assert b"null" in L.tx_last_error()
    assert L.tx_dedisp_run(None) == _lib.TX_EINVAL
    assert L.tx_boxcar_run(None, C.c_float(7.0)) == _lib.TX_EINVAL
    assert L.tx_plan_destroy(None) == _lib.TX_OK


def test_frontend_argument_checks_on_host():
    """tx_frontend_*: the arguments are checked before anything reaches the
    device (dummy pointers here are never dereferenced)."""
    from transientx_amd import _lib

    L = _lib.lib()
    dummy = C.c_void_p(64)
    assert L.tx_frontend_stats_work_bytes(4096, 1) == 4096 * 16
    assert L.tx_frontend_stats_work_bytes(4096, 4) == 1024 * 16
    assert L.tx_frontend_stats_work_bytes(3, 4) == 0 and L.tx_frontend_stats_work_bytes(10, 0) == 0
    assert L.tx_beam_set_frontend(None, 128, 1, 2, dummy, dummy) == _lib.TX_EINVAL
    assert L.tx_beam_set_frontend_in_load(None, 1) == _lib.TX_EINVAL
    assert L.tx_session_begin_host(None, None, 16) == _lib.TX_EINVAL
    # null pointers, td*fd > 256, fewer channels than fd
    assert L.tx_frontend_u8(None, None, 100, 64, 1, 1, dummy, dummy, None, dummy) == _lib.TX_EINVAL
    assert L.tx_frontend_u8(None, dummy, 100, 64 * 257, 1, 257, dummy, dummy, None, dummy) == _lib.TX_EINVAL
    assert L.tx_frontend_u8(None, dummy, 100, 64, 16, 17, dummy, dummy, None, dummy) == _lib.TX_EINVAL
    assert L.tx_frontend_u8(None, dummy, 100, 3, 1, 4, dummy, dummy, None, dummy) == _lib.TX_EINVAL
    # fewer raw samples than td: nothing to write
    assert L.tx_frontend_u8(None, dummy, 1, 64, 2, 1, dummy, dummy, None, dummy) == _lib.TX_OK
    nb = L.tx_frontend_stats_work_bytes(64, 2)
    assert L.tx_frontend_stats(None, None, 100, 64, 1, 2, dummy, dummy, dummy, nb) == _lib.TX_EINVAL
    assert L.tx_frontend_stats(None, dummy, 1, 64, 2, 2, dummy, dummy, dummy, nb) == _lib.TX_EINVAL
    assert L.tx_frontend_stats(None, dummy, 100, 64, 1, 2, dummy, dummy, dummy, nb - 1) == _lib.TX_EINVAL
    assert b"work" in L.tx_last_error()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from transientx_amd import _lib

    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIBPATH", tmp_path / "absent.so")
    with pytest.raises(RuntimeError, match="missing"):
        _lib.lib()


def test_cpp_dropin_header_compiles_and_links(tmp_path):
    """include/transientx_gpu.hpp compiles as plain C++17 and its test
    program links against libtxgpu.so (usage exit without a GPU call)."""
    import shutil
    import subprocess

    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    src = ROOT / "tests" / "cpp" / "dropin_search.cpp"
    out = tmp_path / "dropin"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", str(ROOT / "include"), str(src),
                    "-o", str(out), "-L", str(ROOT / "transientx_amd"), "-ltxgpu",
                    f"-Wl,-rpath,{ROOT / 'transientx_amd'}", "-Wl,-rpath-link,/opt/rocm/lib"],
                   check=True)
    r = subprocess.run([str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def test_cpp_dropin_takes_every_reference_call(tmp_path):
    """tests/cpp/dropin_calls.cpp repeats every expression TransientX's own
    callers write on SubbandDedispersion and Boxcar (SinglePulse::prepare /
    run with --savetim, Boxcar::match's reads, CandPlot, Cluster, the
    driver's flush and makeinf / modifynblock): it compiles against the
    drop-in at the reference's -std=c++11 with -Wall -Werror and links."""
    import shutil
    import subprocess

    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    src = ROOT / "tests" / "cpp" / "dropin_calls.cpp"
    out = tmp_path / "dropin_calls"
    subprocess.run(["g++", "-std=c++11", "-Wall", "-Wextra", "-Werror", "-I", str(ROOT / "include"), str(src),
                    "-o", str(out), "-L", str(ROOT / "transientx_amd"), "-ltxgpu",
                    f"-Wl,-rpath,{ROOT / 'transientx_amd'}", "-Wl,-rpath-link,/opt/rocm/lib"],
                   check=True)
    r = subprocess.run([str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "usage" in r.stdout



def test_launch_sink_bookkeeping_under_asan(tmp_path):
    """The graph pipelines' launch records (transientx_amd/csrc/launch_sink.h):
    argument widths, alignment, KNode copies and sink growth, in a CPU
    program built with AddressSanitizer and UndefinedBehaviorSanitizer."""
    import shutil
    import subprocess

    if not shutil.which("g++") or not Path("/opt/rocm/include/hip/hip_runtime_api.h").exists():
        pytest.skip("g++ or the HIP runtime headers not available")
    src = ROOT / "tests" / "cpp" / "launch_sink_test.cpp"
    out = tmp_path / "launch_sink_test"
    subprocess.run(["g++", "-std=c++17", "-g", "-O1", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                    "-fno-sanitize-recover=all", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", str(src),
                    "-o", str(out), "-L", "/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"], check=True)
    r = subprocess.run([str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "launch_sink ok" in r.stdout


def _sanitized_run(tmp_path, compiler, flags, srcs, name, args=()):
    import shutil
    import subprocess

    if not shutil.which(compiler):
        pytest.skip(f"{compiler} not available")
    out = tmp_path / name
    subprocess.run([compiler, *flags, "-g", "-O1", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                    "-fno-sanitize-recover=all", *map(str, srcs), "-o", str(out), "-lm"], check=True)
    r = subprocess.run([str(out), *map(str, args)], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    assert r.returncode == 0, r.stderr[-2000:]
    assert f"{name} ok" in r.stdout


def test_host_code_under_sanitizers(tmp_path):
    """csrc/host.cpp, plan.cpp and sigproc.cpp on their own with
    AddressSanitizer and UndefinedBehaviorSanitizer: ddplan/JSON parsing of
    the reference's examples and of broken inputs, ndump, clustering of
    random point sets with small caps, kadane, the SIGPROC reader
    (tests/cpp/host_sanitize.cpp)."""
    csrc = ROOT / "transientx_amd" / "csrc"
    _sanitized_run(tmp_path, "g++", ["-std=c++17", "-I", ROOT / "include", "-I", csrc],
                   [ROOT / "tests" / "cpp" / "host_sanitize.cpp", csrc / "host.cpp", csrc / "plan.cpp",
                    csrc / "sigproc.cpp"], "host_sanitize",
                   [ROOT / "tests" / "golden" / "ddplan_example.txt",
                    ROOT / "tests" / "golden" / "transientx_config_example.json"])


def test_oracle_under_sanitizers(tmp_path):
    """oracle/txoracle.c with AddressSanitizer and UndefinedBehaviorSanitizer
    over small and degenerate cases of every family of its functions
    (tests/cpp/oracle_sanitize.c)."""
    _sanitized_run(tmp_path, "gcc", ["-std=c11", "-ffp-contract=off", "-w"],
                   [ROOT / "tests" / "cpp" / "oracle_sanitize.c"], "oracle_sanitize")
