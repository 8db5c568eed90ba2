"""Oracle <-> device tensor conversion for the parity tests."""
import numpy as np
import torch

from oracle.bsten import BSTensor, Index
from ultradmrg_amd.tensor import Axis, BlockTensor


def to_axis(ix: Index) -> Axis:
    return Axis(tuple(q for q, _ in ix.sectors), tuple(d for _, d in ix.sectors), ix.dir)


def to_index(ax: Axis) -> Index:
    return Index(tuple((tuple(q), int(d)) for q, d in zip(ax.qns, ax.degs)), ax.dir)


def to_device(t: BSTensor, device="cuda") -> BlockTensor:
    return BlockTensor.from_blocks([to_axis(ix) for ix in t.indexes], t.blocks, device=device,
                                   dtype=torch.complex128 if t.dtype.kind == "c" else torch.float64)


def to_oracle(t: BlockTensor, like: BSTensor) -> BSTensor:
    out = BSTensor(like.indexes, np.complex128 if t.is_complex else np.float64)
    out.blocks = {k: v.copy() for k, v in t.blocks().items()}
    return out


def from_device(t: BlockTensor) -> BSTensor:
    """The oracle tensor with ``t``'s indexes and blocks (host copy)."""
    out = BSTensor([to_index(ax) for ax in t.axes], np.complex128 if t.is_complex else np.float64)
    out.blocks = {k: v.copy() for k, v in t.blocks().items()}
    return out
