"""Helpers of the GPU tests."""
import os
import sys

import numpy as np


def announce(*what):
    """Names the case on stdout before a GPU step (flushed)."""
    print("[gpu case]", *what, flush=True)
