"""Helpers for batch tests: concatenating CSRs into one batch with global ids."""
import numpy as np


def concat_csr(graphs):
    """[(beg, adj), ...] -> (beg, adj, vbase) with graph g on [vbase[g], vbase[g+1])."""
    begs, adjs, vbase = [np.zeros(1, np.int64)], [], [0]
    e0 = 0
    for beg, adj in graphs:
        beg = np.asarray(beg, np.int64)
        adj = np.asarray(adj, np.int32)
        v0 = vbase[-1]
        begs.append(beg[1:] + e0)
        adjs.append(adj + np.int32(v0))
        e0 += len(adj)
        vbase.append(v0 + len(beg) - 1)
    adj = np.concatenate(adjs) if adjs else np.zeros(0, np.int32)
    return np.concatenate(begs), adj.astype(np.int32), np.array(vbase, np.int64)
