"""The C-ABI library without a GPU: it loads, exports every declared symbol and nothing else,
and its argument checks return the reference's statuses (include/spektralab.h:181-386,
src/abi/abi.cpp:40-92)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

import spektralab_engine_amd as spk

ROOT = Path(__file__).resolve().parents[1]


def declared(header):
    text = (ROOT / "include" / header).read_text()
    return set(re.findall(r"SPK_API\s+[\w\s\*]+?\b(spkx?_\w+)\s*\(", text))


def test_header_declares_the_reference_entry_points():
    assert declared("spektralab.h") == set(spk.REFERENCE_SYMBOLS)
    assert declared("spektralab_ext.h") == set(spk.EXTENSION_SYMBOLS) | {"spkx_plan"}


def test_library_exports_exactly_the_declared_symbols(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", str(spk.LIB_PATH)], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.split() and l.split()[1] in "TWiV"}
    want = declared("spektralab.h") | declared("spektralab_ext.h")
    assert want <= exported
    assert {s for s in exported if not s.startswith("spk")} == set()


def test_struct_layouts():
    assert C.sizeof(spk.Image) == 24 and C.sizeof(spk.DeviceImage) == 24 and C.sizeof(spk.Result) == 88
    assert spk.Result.elapsed_ms.offset == 32 and spk.Result.progress_id.offset == 48


def test_build_info(lib):
    assert re.fullmatch(rb"spektralab-native 1\.0\.0 \(\d{4}-\d{2}-\d{2}, math=safe\)", lib.spk_build_info())


def test_null_arguments(lib):
    res = spk.Result()
    assert lib.spk_render(None, b"full", C.byref(res)) == -1
    assert lib.spk_reprint(None, b"full", None) == -1
    assert lib.spk_set_params(None, b"{}", None) == -1
    assert lib.spk_get_params(None, None) == -1
    assert lib.spk_progress(None, None, None) == -1
    assert lib.spk_warm_up(None, None, None, None) == -1
    assert lib.spk_capabilities(None) == b"" and lib.spk_params_schema(None) == b""
    assert lib.spk_print_lut_catalog(None) == b"{}"
    assert lib.spk_engine_trim(None) == 0
    lib.spk_engine_destroy(None)
    lib.spk_session_release(None)
    lib.spk_result_free(None)
    lib.spk_string_free(None)
    lib.spk_cancel(None, None)
    assert lib.spk_engine_create(None, None) is None
    assert b"NULL" in lib.spk_last_error(None)


def test_engine_refuses_a_missing_folder(lib, tmp_path):
    assert lib.spk_engine_create(str(tmp_path / "nope").encode(), None) is None
    assert lib.spk_last_error(None) != b""


def test_corrupt_blob_is_an_io_error(lib, tmp_path):
    (tmp_path / "spektralab_constants.bin").write_bytes(b"SPKX" + b"\0" * 12)
    assert lib.spk_engine_create(str(tmp_path).encode(), None) is None
    msg = lib.spk_last_error(None)
    assert b"magic" in msg or b"HIP" in msg


def test_plan_refuses_bad_delta(resources):
    with pytest.raises(spk.SpkError, match="unknown parameter"):
        spk.host_plan(resources, {"no_such_field": 1}, 8, 8)
    with pytest.raises(spk.SpkError, match="outside"):
        spk.host_plan(resources, {"print_exposure": 99.0}, 8, 8)
