"""hipGraph capture of batched solves.

The fused solve is a fixed launch sequence with no host synchronisation (the
SQP and PCG stop rules are evaluated on the device), so a whole step --
several solves of a ragged mixed batch, their backward passes and whatever
the caller computes around them -- can be captured once and replayed as one
graph launch.  ``GraphedStep(fn)`` warms ``fn`` up on a side stream,
captures one call with ``torch.cuda.graph`` and ``replay()`` re-runs it on the
tensors captured (static inputs: write new data into them in place).

Requirements for ``fn``: every tensor it reads is on the device before the
capture (problem parameters, guesses, weights with requires_grad), and it
does not read values back to the host.  The library's fused path allocates
nothing and records no memset nodes; torch fills zeros with kernels.
"""
import torch


class GraphedStep:
    def __init__(self, fn, warmup: int = 2):
        self.fn = fn
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(warmup):
                fn()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.out = fn()

    def replay(self):
        """One launch of the captured step; returns the captured outputs
        (overwritten in place by every replay)."""
        self.graph.replay()
        return self.out
