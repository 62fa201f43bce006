"""Locations of the in-tree build products (libraries, code objects, executables)."""
from __future__ import annotations

import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_DIR = os.path.join(PKG_DIR, "lib")
KERNEL_DIR = os.path.join(LIB_DIR, "kernels")
BIN_DIR = os.path.join(LIB_DIR, "bin")


def lib_path(name: str) -> str:
    return os.path.join(LIB_DIR, name)


def kernel_path(name: str) -> str:
    return os.path.join(KERNEL_DIR, name)


def bin_path(name: str) -> str:
    return os.path.join(BIN_DIR, name)


def include_dir() -> str:
    return os.path.join(REPO_DIR, "include")
