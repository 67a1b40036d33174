"""include/pocket_sdr_amd.h coexists with the reference header.

The reference's src/pocket_sdr.h includes <libusb-1.0/libusb.h>, which this
image lacks, so the header itself is not compiled here.  Instead the test reads
it at run time (when the reference tree is mounted) and extracts what a
receiver translation unit that includes both headers would see of the
symbols this library also declares: the typedefs of the hot path's types,
every prototype of the `// sdr_func.c` block whose types are all known, and
the reference's own sdr_arch_combine (src/pocket_sdr.h:522).  Those lines,
then include/pocket_sdr_amd.h, are compiled with gcc (C) and g++ (C++): any
prototype of part 1 that differs from the reference's is a "conflicting
types" error, and so is an additive entry point that reuses a reference name
with another signature.  No GPU is used.
"""
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_H = "/root/reference/src/pocket_sdr.h"

pytestmark = pytest.mark.skipif(not os.path.exists(REF_H), reason="reference tree not mounted")

KNOWN = {"void", "int", "float", "double", "char", "const", "sdr_cpx_t", "sdr_cpx16_t", "sdr_cpx8_t",
         "sdr_buff_t", "int8_t", "uint8_t", "int16_t", "uint16_t", "int32_t", "uint32_t", "size_t",
         "sdr_arch_t", "sdr_rcv_t", "unsigned"}


def extract():
    src = open(REF_H).read()
    guard = re.search(r"#ifndef\s+(\w+)\s*\n#define\s+\1", src).group(1)
    out = ["#include <stdint.h>", "#include <stddef.h>", "#define %s" % guard]
    # the hot path's types, as the reference defines them
    for name in ("sdr_cpx8_t", "sdr_cpx16_t", "sdr_cpx_t"):
        m = re.search(r"^typedef.*\b%s\b[^;\n]*;" % name, src, re.M)
        out.append(m.group(0))
    m = re.search(r"^typedef struct \{[^}]*\} sdr_buff_t;", src, re.M | re.S)
    out.append(m.group(0))
    # receiver structs appear only by pointer in the prototypes kept below
    out += ["typedef struct ref_arch_tag sdr_arch_t;", "typedef struct ref_rcv_tag sdr_rcv_t;"]
    # prototypes of the sdr_func.c block
    blk = src[src.index("// sdr_func.c"):src.index("// sdr_code.c")]
    protos = [p.strip() + ";" for p in " ".join(blk.splitlines()[1:]).split(";") if p.strip()]
    kept = []
    for p in protos:
        words = set(re.findall(r"[A-Za-z_]\w*", re.sub(r"\bsdr_\w+\s*\(", "(", p)))
        types = {w for w in words if w.endswith("_t") or w in KNOWN}
        if types <= KNOWN:
            kept.append(p)
    m = re.search(r"^void sdr_arch_combine\([^;]*;", src, re.M | re.S)
    kept.append(" ".join(m.group(0).split()))
    return out, kept


def test_extracted_prototypes_cover_the_hot_path():
    _, kept = extract()
    names = {re.search(r"\b(sdr_\w+)\s*\(", p).group(1) for p in kept}
    for n in ("sdr_func_init", "sdr_cpx_malloc", "sdr_cpx_free", "sdr_cpx_abs", "sdr_cpx_mul",
              "sdr_cpx_fft", "sdr_buff_new", "sdr_buff_free", "sdr_search_code", "sdr_corr_max",
              "sdr_fine_dop", "sdr_dop_bins", "sdr_corr_std", "sdr_corr_std_cpx_code",
              "sdr_corr_std_cpx", "sdr_corr_fft", "sdr_corr_fft_cpx", "sdr_mix_carr",
              "sdr_arch_combine"):
        assert n in names, n


@pytest.mark.parametrize("compiler", [("gcc", "c", "-std=c99"), ("g++", "c++", "-std=c++17")])
def test_reference_prototypes_and_header_compile_together(compiler):
    cc, lang, std = compiler
    head, kept = extract()
    code = "\n".join(head)
    if lang == "c++":
        code += '\nextern "C" {\n' + "\n".join(kept) + "\n}\n"
    else:
        code += "\n" + "\n".join(kept) + "\n"
    code += '#include "pocket_sdr_amd.h"\nint main(void) { return 0; }\n'
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "both." + ("cpp" if lang == "c++" else "c"))
        open(f, "w").write(code)
        r = subprocess.run([cc, std, "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), f],
                           capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
