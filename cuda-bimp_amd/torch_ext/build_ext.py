"""Builds bimp_amd/bimp_torch*.so (the libtorch autograd Function) with g++ against the
installed torch, linked to the in-tree libbimp_hip.so through an $ORIGIN rpath."""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "bimp_amd")
INC = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include")


def build():
    import torch
    from torch.utils import cpp_extension as ce
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(PKG, "bimp_torch" + suffix)
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = ["g++", "-shared", "-fPIC", "-O2", "-std=c++17", "-w",
           f"-D_GLIBCXX_USE_CXX11_ABI={abi}", "-DTORCH_EXTENSION_NAME=bimp_torch",
           "-DTORCH_API_INCLUDE_EXTENSION_H", f"-I{INC}",
           f"-I{sysconfig.get_paths()['include']}"]
    cmd += [f"-I{p}" for p in ce.include_paths()]
    cmd += [os.path.join(HERE, "bimp_autograd.cpp"), "-o", out + ".tmp"]
    cmd += [f"-L{p}" for p in ce.library_paths()]
    cmd += ["-lc10", "-ltorch", "-ltorch_cpu", "-ltorch_python", f"-L{PKG}", "-lbimp_hip",
            "-Wl,-rpath,$ORIGIN"]
    subprocess.check_call(cmd)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build())
    sys.exit(0)
