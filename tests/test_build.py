"""The build entry points work from a tree that has no built library: the
package imports without libglp.so (so `python -m glp_amd.build` and
__graft_entry__.build() can run), and the first call into the library then
fails loudly instead of falling back to anything."""
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_package_and_build_module_import_without_library(tmp_path):
    pkg = tmp_path / "glp_amd"
    pkg.mkdir()
    for f in ("__init__.py", "build.py"):
        shutil.copy(os.path.join(ROOT, "glp_amd", f), pkg / f)
    assert not (pkg / "libglp.so").exists()
    code = (
        "import glp_amd\n"
        "from glp_amd import build\n"
        "import glp_amd.build as b2\n"
        "assert build is b2 and callable(build.build)\n"
        "assert build.LIB == glp_amd.LIB_PATH\n"
        "try:\n"
        "    glp_amd.version()\n"
        "except ImportError as e:\n"
        "    assert 'libglp.so not built' in str(e)\n"
        "    print('loud')\n"
        "else:\n"
        "    raise SystemExit('a call without libglp.so did not raise')\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=str(tmp_path), capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "loud"
