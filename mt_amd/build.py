"""Builds the engine and its C++ host in-tree.

  libmt_engine.so  the gfx950 engine (hipcc); every HIP source is compiled
                   with packed-FP32 instruction selection disabled
                   (-Xclang -target-feature -Xclang -packed-fp32-ops)
  libmt_host.so    the C++ host: configuration files, PDB/DCD/XYZ files,
                   run setup, updater and step loop (g++, links the engine)
  mt_host          the command-line driver on top of both
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmt_engine.so")
SOURCES = ["engine.hip", "forces.hip", "lists.hip", "cells.hip", "tea.hip", "mlp.hip", "resident.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
HOST = os.path.join(HERE, "host")
HOST_SOURCES = ["config.cpp", "pdbio.cpp", "setup.cpp", "updater.cpp", "capi.cpp"]
HOST_LIB = os.path.join(HERE, "libmt_host.so")
DRIVER = os.path.join(HERE, "mt_host")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
CXXFLAGS = ["-O2", "-std=c++17", "-Wall", "-Wextra", "-Wno-unused-parameter", "-I" + INCLUDE]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics",
         "-fno-hip-fp32-correctly-rounded-divide-sqrt"]


def _run(cmd, what, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build of {what} failed")


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, *FLAGS, "-o", tmp, *srcs], "libmt_engine.so", verbose)
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


def build_host(verbose: bool = False) -> str:
    """g++ build of the host library and driver against libmt_engine.so."""
    link = ["-L" + HERE, "-lmt_engine", "-pthread", "-Wl,-rpath,$ORIGIN"]
    tmp = HOST_LIB + ".tmp"
    _run([CXX, *CXXFLAGS, "-fPIC", "-shared", "-o", tmp, *[os.path.join(HOST, s) for s in HOST_SOURCES], *link],
         "libmt_host.so", verbose)
    os.replace(tmp, HOST_LIB)
    tmp = DRIVER + ".tmp"
    _run([CXX, *CXXFLAGS, "-o", tmp, os.path.join(HOST, "driver.cpp"), "-L" + HERE, "-lmt_host", *link],
         "mt_host", verbose)
    os.replace(tmp, DRIVER)
    return HOST_LIB


if __name__ == "__main__":
    print(build(verbose=True))
