"""CPU restatement of the matrix-block <-> buffer copies around the grid
calls (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ -- never by the product package.

  pack_buffer    rs_pack_buffer, src/task_list_methods.F:2597-2630
                 (via rs_copy_to_buffer :2561-2572): each atom pair's block,
                 flattened column-major (RESHAPE), lands at its pair offset.
  unpack_buffer  rs_unpack_buffer, src/task_list_methods.F:2636-2674
                 (via rs_copy_to_matrices :2578-2589): each block is
                 incremented by its slice of the buffer.

Pinned in tests/test_oracle.py by the round trip pack -> unpack into zero
blocks (returns the blocks) and by the task-list layout (packing the blocks
of a generated density matrix reproduces the buffer the generator wrote).
"""
import numpy as np


def pack_buffer(blocks, pair_offsets, size, buf=None):
    """blocks: 1-D arrays (column-major block data); returns the buffer."""
    out = np.zeros(size) if buf is None else buf
    for blk, off in zip(blocks, pair_offsets):
        flat = np.asarray(blk, dtype=np.float64).reshape(-1)
        out[int(off):int(off) + flat.size] = flat
    return out


def unpack_buffer(blocks, pair_offsets, buf):
    """blocks[i] += buf[pair_offsets[i] : + blocks[i].size], in place."""
    for blk, off in zip(blocks, pair_offsets):
        blk += buf[int(off):int(off) + blk.size]
    return blocks
