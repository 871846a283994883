"""The ctypes mirrors in louvain_amd.py have the layout of the C structs in
include/louvain.h: size and the offset of every field, as gcc lays them out
(no GPU needed)."""
import ctypes
import os
import shutil
import subprocess

import pytest

import louvain_amd as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PAIRS = [
    ("lv_options", L.Options),
    ("lv_stats", L.Stats),
    ("lv_kernel_time", L.KernelTime),
    ("lv_frontier", L.Frontier),
    ("lv_train_options", L.TrainOptions),
    ("lv_batch_result", L.BatchResult),
]


def _c_layout(tmp_path):
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "louvain.h"', "int main(void) {"]
    for cname, py in PAIRS:
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'    printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    exe = tmp_path / "layout"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {tuple(line.split()[:2]): int(line.split()[2]) for line in out.splitlines()}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_ctypes_structs_match_header(tmp_path):
    c = _c_layout(tmp_path)
    for cname, py in PAIRS:
        assert ctypes.sizeof(py) == c[(cname, "size")], cname
        for fname, _ in py._fields_:
            assert getattr(py, fname).offset == c[(cname, fname)], (cname, fname)
