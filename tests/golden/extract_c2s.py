"""Extract the reference's committed c2s golden tables as data.

Reads the two mpmath-generated coefficient tables that the reference keeps in
its tests (crates/integral/tests/c2s_independent_reference.rs C2S_SIGNED_REF,
signed, l <= 6; crates/integral/tests/c2s_reference.rs C2S_REF, rows
sign-fixed) and writes their numbers to c2s_reference.json.  Only the numeric
rows are kept; run in a checkout that has the reference at /root/reference.
"""
import json
import os
import re

REF = "/root/reference/crates/integral/tests"
HERE = os.path.dirname(os.path.abspath(__file__))


def table(path, name):
    src = open(path).read()
    start = src.index("pub const %s" % name)
    body = src[src.index("= &[", start) + 4:]
    depth, end = 1, 0
    for k, ch in enumerate(body):
        if ch == "[":
            depth += 1
        elif ch == "]":
            depth -= 1
            if depth == 0:
                end = k
                break
    body = re.sub(r"//[^\n]*", "", body[:end])
    rows = re.findall(r"&\[([^\]]*)\]", body)
    return [[float(x) for x in r.replace("\n", " ").split(",") if x.strip()] for r in rows]


def main():
    out = {
        "signed": table(os.path.join(REF, "c2s_independent_reference.rs"), "C2S_SIGNED_REF"),
        "sign_fixed": table(os.path.join(REF, "c2s_reference.rs"), "C2S_REF"),
        "source": "crates/integral/tests/c2s_independent_reference.rs (signed), "
                  "crates/integral/tests/c2s_reference.rs (sign-fixed); rows l=0..6 in m_order",
    }
    with open(os.path.join(HERE, "c2s_reference.json"), "w") as f:
        json.dump(out, f, indent=0)
    print(len(out["signed"]), len(out["sign_fixed"]))


if __name__ == "__main__":
    main()
