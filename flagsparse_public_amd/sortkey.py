"""Stable orders of nonnegative integer keys with a known bound.

A device radix sort makes one pass per key byte, so keys that fit 16 or 32
bits are sorted in that type; a stable sort orders equal keys the same
whatever their type, so the order does not depend on the narrowing.
"""

from __future__ import annotations

import torch


def narrow_key(key: torch.Tensor, bound: int) -> torch.Tensor:
    """``key`` (values in [0, bound)) in the narrowest signed integer type
    that holds them."""
    if bound <= 2**15:
        return key.to(torch.int16)
    if bound <= 2**31:
        return key.to(torch.int32)
    return key


def stable_order(key: torch.Tensor, bound: int) -> torch.Tensor:
    """Indices that sort ``key`` (values in [0, bound)) stably."""
    return torch.argsort(narrow_key(key, bound), stable=True)


def stable_sort(key: torch.Tensor, bound: int):
    """(sorted keys in the narrowed type, indices) of a stable sort of
    ``key`` (values in [0, bound)): the sorted keys come with the sort, not
    from a gather afterwards."""
    return torch.sort(narrow_key(key, bound), stable=True)
