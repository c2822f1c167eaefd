"""Compiles the C++ test programs against include/lsmps3d_amd.hpp and
lsmps-3d_amd/liblsmps3d_amd.so into tests/cpp/_build/:
module_step (the module API), run_simulation (the application loop) and the
three diagnostic drivers (hydrostatic_surface_diagnostics, whose pressure
operator is a HIP kernel of its own, complex_surface_diagnostics and
pipe_flow_diagnostics)."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIBDIR = os.path.join(ROOT, "lsmps-3d_amd")
OUT = os.path.join(HERE, "_build", "module_step")


PROGRAMS = ("module_step.cpp", "run_simulation.cpp", "hydrostatic_surface_diagnostics.hip",
            "complex_surface_diagnostics.cpp", "pipe_flow_diagnostics.cpp")


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    for source in PROGRAMS:
        name = os.path.splitext(source)[0]
        out = os.path.join(HERE, "_build", name)
        cmd = [hipcc, "-std=c++17", "-O2", "--offload-arch=gfx950", "-Wall", "-Wextra",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               f"-I{os.path.join(ROOT, 'include')}", os.path.join(HERE, source),
               f"-L{LIBDIR}", "-llsmps3d_amd", f"-Wl,-rpath,{LIBDIR}",
               "-Wl,-rpath,$ORIGIN/../../../lsmps-3d_amd", "-o", out + ".tmp"]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed for {source}:\n{res.stdout}\n{res.stderr}")
        os.replace(out + ".tmp", out)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
