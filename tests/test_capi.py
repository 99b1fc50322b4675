"""CPU-side checks of the C-ABI library: it loads, exports every entry point
declared in include/tremor_amd/tmx.h, and rejects use without a device.
No compute call is made here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tremor_amd", "tmx.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(tmx_[a-z_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from tremor_amd import capi
    if not os.path.exists(capi.LIB_PATH):
        from tremor_amd import build
        build.build()
    return ctypes.CDLL(capi.LIB_PATH)


def test_exports_every_declared_symbol(lib):
    names = _declared()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), f"{n} declared in tmx.h but not exported"


def test_binding_covers_header():
    from tremor_amd import capi
    assert sorted(capi.EXPORTS) == _declared()


def test_cfg_default_and_null_handling(lib):
    from tremor_amd import capi
    L = capi.lib()
    cfg = capi.Cfg()
    assert L.tmx_cfg_default(ctypes.byref(cfg)) == 0
    assert cfg.precision == capi.PREC_BF16X3 and cfg.device == 0
    assert L.tmx_last_error(None) == b"null context"
    assert L.tmx_sync(None) != 0
    assert L.tmx_waveform_length(None) == 0


def test_invalid_precision_rejected_before_device(lib):
    from tremor_amd import capi
    L = capi.lib()
    cfg = capi.Cfg()
    L.tmx_cfg_default(ctypes.byref(cfg))
    cfg.precision = 7
    h = ctypes.c_void_p()
    assert L.tmx_create(ctypes.byref(h), ctypes.byref(cfg)) != 0
    assert not h.value


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes mirror of a C-ABI struct (tmx_cfg, tmx_stats, tmx_hits,
    tmx_server_cfg, tmx_request, tmx_result) has the
    header's field offsets and sizes (compiled with the host C compiler)."""
    import shutil
    import subprocess
    from tremor_amd import capi
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    structs = {"tmx_cfg": capi.Cfg, "tmx_stats": capi.Stats, "tmx_hits": capi.Hits, "tmx_server_cfg": capi.ServerCfg,
               "tmx_request": capi.Request, "tmx_result": capi.Result}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "tremor_amd/tmx.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(line.split()[:2]): int(line.split()[2]) for line in out if line.strip()}
    for cname, py in structs.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
