"""ABI layout of a copy of mr_aha.h: every struct field's offset and size,
every struct's size and alignment, every enumerator and constant of the
public header. TEST INFRASTRUCTURE (oracle side).

Built against the reference's own header (include/mr_aha/mr_aha.h under the
reference checkout; it compiles on its own with gcc) into oracle/_ref/, and
against this repository's header; tests/test_abi.py compares the two line
for line. The field lists come from the binding's ctypes mirrors: a field
missing from either header fails to compile, a field missing from a mirror
changes the struct size.
"""

from __future__ import annotations

import re
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent
REF_INCLUDE = Path("/root/reference/include")
REF_OUT = HERE / "_ref" / "abi_ref.txt"

STRUCTS = {"mr_config": "MRConfig", "mr_op_inputs": "MROpInputs",
           "mr_op_options": "MROpOptions", "mr_op_info_t": "MROpInfo"}


def _constants(header: Path):
    """Enumerators and MR_ macros of a header (names only)."""
    text = header.read_text()
    names = set(re.findall(r"^\s*(MR_[A-Z0-9_]+)\s*(?:=|,|$)", text, re.M))
    names |= set(re.findall(r"^#define\s+(MR_[A-Z0-9_]+)\s+\S", text, re.M))
    names.discard("MR_AHA_MR_AHA_H")
    return sorted(n for n in names if not n.endswith("_H"))


def program() -> str:
    """C source printing the layout (field lists from mr_aha_amd._lib)."""
    # field names read from the mirrors' source (the module itself loads the
    # library at import)
    src = (REPO / "mr_aha_amd" / "_lib.py").read_text()
    fields = {}
    for cname, pyname in STRUCTS.items():
        m = re.search(r"class %s\(ctypes\.Structure\):\s*_fields_ = \[(.*?)\]\n\n" % pyname, src,
                      re.S)
        fields[cname] = re.findall(r'\(\s*"(\w+)"', m.group(1))
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "mr_aha/mr_aha.h"',
             "#define F(T, f) printf(#T \".\" #f \" %zu %zu\\n\", offsetof(T, f), "
             "sizeof(((T *)0)->f));",
             "#define C(n) printf(#n \" %lld\\n\", (long long)(n));",
             "int main(void) {"]
    for cname, fl in fields.items():
        lines.append(f'  printf("{cname} %zu %zu\\n", sizeof({cname}), _Alignof({cname}));')
        lines += [f"  F({cname}, {f})" for f in fl]
    lines.append('  printf("mr_complex %zu %zu\\n", sizeof(mr_complex), _Alignof(mr_complex));')
    lines += [f"  C({n})" for n in _constants(REPO / "include" / "mr_aha" / "mr_aha.h")]
    lines += ["  return 0;", "}"]
    return "\n".join(lines) + "\n"


def run(include_dir: Path, work: Path) -> str:
    work.mkdir(parents=True, exist_ok=True)
    src = work / "abi_layout.c"
    exe = work / "abi_layout"
    src.write_text(program())
    subprocess.run(["gcc", "-std=c11", f"-I{include_dir}", str(src), "-o", str(exe)], check=True)
    return subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout


def reference_layout() -> str | None:
    """The reference header's layout: built from the reference checkout when it
    is present (written to oracle/_ref/abi_ref.txt), else that file, else None."""
    if (REF_INCLUDE / "mr_aha" / "mr_aha.h").exists():
        out = run(REF_INCLUDE, HERE / "_ref")
        REF_OUT.write_text(out)
        return out
    return REF_OUT.read_text() if REF_OUT.exists() else None


if __name__ == "__main__":
    print(reference_layout())
