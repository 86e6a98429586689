"""The ctypes mirrors of the C-ABI's config structs (cbrkc.Config for
cb_pop_config, cbrkc.NetConfig for cb_net_config) have the C header's
layout: every field at the offset gcc gives it, and the same total size, so
the *_config_default memsets and the create calls' reads see the fields the
Python side sets."""
import ctypes
import os
import subprocess

import pytest

import cbrkc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("cname,pyclass", [("cb_pop_config", cbrkc.Config), ("cb_net_config", cbrkc.NetConfig)])
def test_config_layout_matches_the_header(tmp_path, cname, pyclass):
    names = [f[0] for f in pyclass._fields_]
    src = tmp_path / "layout.c"
    body = "\n".join(f'  printf("{n} %zu\\n", offsetof({cname}, {n}));' for n in names)
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "cbrkc.h"\nint main(void) {\n' + body +
                   f'\n  printf("SIZE %zu\\n", sizeof({cname}));\n  return 0;\n}}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                 text=True).stdout.splitlines())
    for n in names:
        assert int(out[n]) == getattr(pyclass, n).offset, n
    assert int(out["SIZE"]) == ctypes.sizeof(pyclass)
