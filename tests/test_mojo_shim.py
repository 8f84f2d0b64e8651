"""The Mojo binding module (mojo/numax_hip.mojo) against the C header, without
a Mojo toolchain: it is what mojo/gen_numax_hip_mojo.py writes from the
current header, and every declared entry point has one wrapper taking the
library handle and then the header's parameters in order."""
import os
import re
import subprocess
import sys

ROOT = os.path.join(os.path.dirname(__file__), "..")


def _header_decls():
    sys.path.insert(0, os.path.join(ROOT, "mojo"))
    import gen_numax_hip_mojo as gen
    return gen, gen.declarations(open(gen.HEADER).read())


def test_module_is_generated_from_the_current_header():
    rc = subprocess.run([sys.executable, os.path.join(ROOT, "mojo", "gen_numax_hip_mojo.py"), "--check"]).returncode
    assert rc == 0, "mojo/numax_hip.mojo is stale: run python mojo/gen_numax_hip_mojo.py"


def test_every_entry_point_has_a_wrapper_with_the_headers_parameters():
    gen, decls = _header_decls()
    text = open(os.path.join(ROOT, "mojo", "numax_hip.mojo")).read()
    wrappers = dict(re.findall(r"^fn (nmx_[a-z0-9_]+)\(h: NmxLib(.*?)\)", text, flags=re.M))
    assert sorted(wrappers) == sorted(name for _, name, _ in decls)
    for _, name, params in decls:
        got = [a.split(":")[0].strip() for a in wrappers[name].split(",") if a.strip()]
        assert got == [gen.pname(p) for p, _ in params], name
        # the function type handed to get_function has one Mojo type per C parameter
        sig = re.search(r'get_function\[fn \((.*?)\) -> [A-Za-z\[\]0-9]+\]\("' + re.escape(name) + r'"\)', text).group(1)
        assert len([t for t in re.split(r",\s*(?![^\[]*\])", sig) if t.strip()]) == len(params), name
