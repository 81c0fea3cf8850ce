"""Properties of the gfx950 machine code in libfpa.so, read from the library's own
offload bundles (objcopy + clang-offload-bundler + llvm-objdump of this build).

Every kernel body must be fully inlined: no kernel may call a device function.  Kernels
whose cores were compiled as real calls (pointer arguments into the private segment,
reached through flat loads/stores) are the only ones that ever failed on the device,
so the build keeps them out: the dot-chain kernel is instantiated only for shapes whose
body inlines, and this test fails if any call appears.

No kernel may contain a FLAT (generic-address) load or store either: every LDS access
must be a DS instruction and every HBM access a global one.  A FLAT instruction picks
its aperture (global, LDS, scratch) from the high bits of its base address, without the
instruction's immediate offset; a guarded read of lds[w+1] at w = -1, compiled as
flat_load (lds - 4) offset:4, then leaves the LDS aperture when lds is LDS address 0 and
faults.  Generic pointers appear only when a core is compiled as a real call, which is
how the two kernels that faulted were built.

The element-wise kernels of one op may be split over several compilation units (parts
that build in parallel): every (op, shape) the dispatcher can launch must be compiled
exactly once, in one code object."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "cubigfloat_amd", "libfpa.so")
LLVM = "/opt/rocm/llvm/bin"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"


def code_objects(tmp_path):
    fat = tmp_path / "fatbin.bin"
    subprocess.run(["objcopy", "-O", "binary", "--only-section=.hip_fatbin", LIB, str(fat)], check=True)
    data = fat.read_bytes()
    starts = [m.start() for m in re.finditer(re.escape(MAGIC), data)]
    assert starts, "no offload bundle in libfpa.so"
    out = []
    for i, s in enumerate(starts):
        chunk = data[s:starts[i + 1] if i + 1 < len(starts) else len(data)]
        src = tmp_path / f"bundle{i}.bin"
        dst = tmp_path / f"bundle{i}.co"
        src.write_bytes(chunk)
        subprocess.run([f"{LLVM}/clang-offload-bundler", "--type=o", "--targets=hipv4-amdgcn-amd-amdhsa--gfx950",
                        f"--input={src}", f"--output={dst}", "--unbundle"], check=True)
        out.append(dst)
    return out


@pytest.mark.skipif(not (os.path.exists(LIB) and shutil.which("objcopy") and os.path.exists(f"{LLVM}/llvm-objdump")),
                    reason="library or LLVM tools missing")
def test_no_kernel_calls_a_function_or_uses_flat_addressing(tmp_path):
    calls, flat = {}, {}
    kernels = 0
    for co in code_objects(tmp_path):
        dis = subprocess.run([f"{LLVM}/llvm-objdump", "-d", str(co)], check=True, capture_output=True,
                             text=True).stdout
        cur = None
        for line in dis.splitlines():
            m = re.match(r"^[0-9a-f]+ <(\S+)>:", line)
            if m:
                cur = m.group(1)
                kernels += "fpa_kernel" in cur or "fpa_dot_kernel" in cur
                continue
            if "s_swappc_b64" in line:
                calls[cur] = calls.get(cur, 0) + 1
            if re.search(r"\bflat_(load|store|atomic)", line):
                flat[cur] = flat.get(cur, 0) + 1
    assert kernels > 0
    assert not calls, f"kernels with device-function calls: {sorted(calls)[:8]}"
    assert not flat, f"kernels with FLAT memory instructions: {sorted(flat)[:8]}"


def _shapes():
    txt = open(os.path.join(ROOT, "cubigfloat_amd", "csrc", "fpa_shapes.h")).read()
    shapes = [tuple(int(v) for v in m) for m in re.findall(r"X\((\d+), (\d+), (\d+)\)", txt)]
    mul_max = int(re.search(r"#define FPA_MUL_MAX_K (\d+)", txt).group(1))
    return shapes, mul_max


@pytest.mark.skipif(not (os.path.exists(LIB) and shutil.which("objcopy") and os.path.exists(f"{LLVM}/llvm-readelf")),
                    reason="library or LLVM tools missing")
def test_every_shape_compiled_once(tmp_path):
    seen = {}
    for i, co in enumerate(code_objects(tmp_path)):
        syms = subprocess.run([f"{LLVM}/llvm-readelf", "--dyn-syms", "-C", str(co)], check=True, capture_output=True,
                              text=True).stdout
        for m in set(re.findall(r"fpa::fpa_kernel<(\d+), (\d+), (\d+)>\(fpa_arguments\)", syms)):
            key = tuple(int(v) for v in m)
            assert key not in seen, f"kernel K={key[0]} L={key[1]} op={key[2]} in two code objects"
            seen[key] = i
    shapes, mul_max = _shapes()
    want = {(K, L, op) for op in range(6) for (_, K, L) in shapes if op != 3 or K <= mul_max}
    assert set(seen) == want, (sorted(want - set(seen))[:5], sorted(set(seen) - want)[:5])
