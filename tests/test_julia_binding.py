"""The Julia binding (bondupdategalerkin_amd/julia/libbug.jl) against include/bug.h, statically:
Julia is not installed here, so the checks are that every ``ccall`` names a declared entry point
and that every mirrored struct has the C struct's fields in order with matching types (a shifted
or mistyped field would corrupt every call that passes the struct)."""
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
HEADER = (ROOT / "include" / "bug.h").read_text()
JULIA = (ROOT / "bondupdategalerkin_amd" / "julia" / "libbug.jl").read_text()

_C_TO_JL = {"int32_t": "Int32", "int64_t": "Int64", "uint64_t": "UInt64", "double": "Float64"}
_STRUCTS = {"bug_leg": "BugLeg", "bug_bsm": "BugBsm", "bug_mpo_site": "BugMpoSite",
            "bug_step_opts": "BugStepOpts", "bug_bond_log": "BugBondLog",
            "bug_step_info": "BugStepInfo", "bug_expv_report": "BugExpvReport"}


def _c_fields(name):
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), HEADER, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    out = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        m = re.match(r"(const\s+)?(\w+)\s*(\**)\s*(.*)", decl, re.S)
        base, stars, names = m.group(2), m.group(3), m.group(4)
        for part in names.split(","):
            part = part.strip()
            p2 = stars
            while part.startswith("*"):
                p2 += "*"
                part = part[1:].strip()
            arr = re.match(r"(\w+)\[(\d+)\]", part)
            if arr:
                out.append((arr.group(1), f"NTuple{{{arr.group(2)},{_C_TO_JL[base]}}}"))
            elif p2:
                out.append((part, "Ptr"))
            else:
                out.append((part, _STRUCTS.get(base, _C_TO_JL.get(base, base))))
    return out


def _jl_fields(name):
    body = re.search(r"(?:mutable )?struct %s\n(.*?)\nend" % name, JULIA, re.S).group(1)
    out = []
    for line in body.splitlines():
        m = re.match(r"\s+(\w+)::(\S+)\s*$", line)
        if m:
            typ = m.group(2)
            out.append((m.group(1), "Ptr" if typ.startswith("Ptr{") else typ))
    return out


def test_every_ccall_names_a_declared_entry_point():
    declared = set(re.findall(r"\b(bug_\w+)\s*\(", HEADER))
    called = set(re.findall(r"\(:(bug_\w+), LIBBUG\)", JULIA))
    called |= set(re.findall(r":(bug_two_site_sketch_(?:left|right))\b", JULIA))
    assert called, "no ccall found"
    assert called <= declared, sorted(called - declared)
    # the step, its handles and the operator seams are all bound
    for sym in ("bug_step", "bug_state_create", "bug_mpo_create", "bug_step_opts_default",
                "bug_two_site_sketch_left", "bug_two_site_sketch_right", "bug_expv",
                "bug_svd_select", "bug_apply_one_site", "bug_apply_zero_site",
                "bug_env_push_left", "bug_env_push_right"):
        assert sym in called, sym


def test_mirrored_structs_match_the_header_field_for_field():
    for c, jl in _STRUCTS.items():
        assert _jl_fields(jl) == _c_fields(c), c


def test_status_and_option_constants_match_the_header():
    for name in ("BUG_OK", "BUG_EINVAL", "BUG_ENONFINITE", "BUG_EINVARIANT", "BUG_EDEVICE",
                 "BUG_ENOMEM", "BUG_PREC_FP64", "BUG_PREC_FP32"):
        c = int(re.search(r"#define %s (\d+)" % name, HEADER).group(1))
        j = int(re.search(r"const %s = \w+\((\d+)\)" % name, JULIA).group(1))
        assert c == j, name
    for sym, name in ((":bond", "BUG_CEIL_BOND"), (":maxdim", "BUG_CEIL_MAXDIM"),
                      (":bond_strict", "BUG_CEIL_BOND_STRICT"),
                      (":frame_root", "BUG_CEIL_FRAME_ROOT")):
        c = int(re.search(r"#define %s (\d+)" % name, HEADER).group(1))
        j = int(re.search(r"%s => Int32\((\d+)\)" % sym, JULIA).group(1))
        assert c == j, name
    for bit, name in re.findall(r"\((\d+), :(\w+)\)", JULIA.split("BUG_UNMET_NAMES")[1][:200]):
        c = int(re.search(r"#define BUG_UNMET_%s (\d+)" % name.upper(), HEADER).group(1))
        assert c == int(bit), name
