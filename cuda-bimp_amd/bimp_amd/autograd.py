"""Differentiable BIMP responses for PyTorch.

`BimpResponses(width, height, lambdas)(images, params)` maps a (N, H, W) float32 device tensor
and params = (sigma_ratio, gamma) to the (N, R) response vector (the D maps of every scale,
ascending wavelength, concatenated).  Forward and backward are the library's differentiation
form (bimp_diff_*) behind a libtorch torch::autograd::Function (bimp_torch extension); the
backward recomputes the complex responses and returns gradients for images and params.
"""
import importlib.machinery
import importlib.util
import os
import sysconfig

import numpy as np
import torch

from . import Diff, default_options

_HERE = os.path.dirname(os.path.abspath(__file__))
_EXT = os.path.join(_HERE, "bimp_torch" + sysconfig.get_config_var("EXT_SUFFIX"))


def _load_ext():
    if not os.path.exists(_EXT):
        raise ImportError(f"{_EXT} is missing: build it with __graft_entry__.build() "
                          "(cuda-bimp_amd/torch_ext/build_ext.py)")
    loader = importlib.machinery.ExtensionFileLoader("bimp_torch", _EXT)
    spec = importlib.util.spec_from_file_location("bimp_torch", _EXT, loader=loader)
    mod = importlib.util.module_from_spec(spec)
    loader.exec_module(mod)
    return mod


_ext = _load_ext()


class BimpResponses:
    def __init__(self, width, height, lambdas, options=None):
        self.diff = Diff(width, height, lambdas, options)
        self.width, self.height = int(width), int(height)
        self.response_size = self.diff.response_size

    def __call__(self, images, params):
        stream = torch.cuda.current_stream(images.device).cuda_stream
        return _ext.response(images, params, int(self.diff.d.value), int(stream))

    def close(self):
        self.diff.close()
