"""The ctypes mirror of cb_pop_config (cbrkc.Config) has the C header's
layout: every field at the offset gcc gives it, and the same total size, so
cb_config_default's memset and cb_pop_create's reads see the fields the
Python side sets."""
import os
import subprocess

import cbrkc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_config_layout_matches_the_header(tmp_path):
    names = [f[0] for f in cbrkc.Config._fields_]
    src = tmp_path / "layout.c"
    body = "\n".join(f'  printf("{n} %zu\\n", offsetof(cb_pop_config, {n}));' for n in names)
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "cbrkc.h"\nint main(void) {\n' + body +
                   '\n  printf("SIZE %zu\\n", sizeof(cb_pop_config));\n  return 0;\n}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                 text=True).stdout.splitlines())
    for n in names:
        assert int(out[n]) == getattr(cbrkc.Config, n).offset, n
    assert int(out["SIZE"]) == ctypes_size()


def ctypes_size():
    import ctypes
    return ctypes.sizeof(cbrkc.Config)
