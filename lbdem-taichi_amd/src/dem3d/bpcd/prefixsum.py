"""Exclusive prefix sum (API of src/dem3d/bpcd/prefixsum.py) on device tensors."""
import torch


class PrefixSumExecutor:
    def serial(self, output, input):
        self.parallel_fast(output, input)

    def parallel_fast(self, output, input, cal_total=False):
        """output <- exclusive scan of input (torch tensors on the same device)."""
        n = input.shape[0]
        if n & (n - 1):
            raise ValueError("parallel_fast requires input count = 2**p")
        inc = torch.cumsum(input.to(torch.int64), 0)
        output.copy_((inc - input.to(torch.int64)).to(output.dtype))
        if cal_total:
            return int(inc[-1])
