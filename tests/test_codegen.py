"""Checks on the gfx950 assembly of the search kernels (built here, no GPU).

v_ashr_pk_u8_i32 writes only the low half of its destination on the device,
while the compiler ORs further bytes into the upper half as if it were zero;
the library is built with that instruction disabled and must not contain it.
LDS accesses stay naturally aligned (anchors are read as whole 8-byte words),
and no access goes through flat (generic-address) instructions: every plane
read is a global load and every window read an LDS one."""
import os
import re
import subprocess

from conftest import ROOT

CSRC = os.path.join(ROOT, "tcodec_amd", "csrc")
ASM = os.path.join(ROOT, "tcodec_amd", "lib", "tcme_kernels.s")
ASM_MC = os.path.join(ROOT, "tcodec_amd", "lib", "tcme_mc.s")
ASM_AUX = os.path.join(ROOT, "tcodec_amd", "lib", "tcme_aux.s")
SIZE = {"b32": 4, "b64": 8, "b96": 16, "b128": 16, "u16": 2, "i16": 2, "b16": 2, "u8": 1,
        "i8": 1, "b8": 1}


def assembly() -> str:
    subprocess.run(["make", "-s", "-C", CSRC, "asm"], check=True)
    return open(ASM).read() + open(ASM_MC).read() + open(ASM_AUX).read()


def test_lds_accesses_are_naturally_aligned():
    text = assembly()
    bad = []
    for m in re.finditer(r"\n\s+ds_(read|write)(2?)_(\w+?) [^\n]*?((?:offset\d?:\d+\s*)*)\n", text):
        kind, two, ty, offs = m.groups()
        size = SIZE.get(ty)
        if size is None or two:
            continue  # read2/write2 offsets count elements, not bytes
        for off in re.findall(r"offset:(\d+)", offs):
            if int(off) % size:
                bad.append(m.group(0).strip())
    assert not bad, bad[:5]


def test_no_packed_shift_clip():
    text = assembly()
    assert "v_ashr_pk_" not in text


def test_no_flat_memory_instructions():
    text = assembly()
    flat = re.findall(r"\n\s+(flat_\w+)", text)
    assert not flat, sorted(set(flat))


def test_every_kernel_present():
    text = assembly()
    for n, wpc in [(8, 1), (16, 1), (32, 1), (32, 2), (32, 4), (64, 2), (64, 4), (64, 8)]:
        assert re.search(rf"me_search_kernelILi{n}ELi{wpc}E", text), (n, wpc)
    for w in (4, 8):
        assert re.search(rf"me_tile_kernelILi{w}E", text), w
    assert "anchor_kernel" in text
    assert "mc_luma_kernel" in text and "mc_chroma_kernel" in text
    assert "gmv_table_kernel" in text and "gmv_walk_kernel" in text and "sad_kernel" in text
