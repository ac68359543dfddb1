"""Learned surrogate of the force-indentation curve: an MLP trained with the libnfx MFMA kernels on
curves from the batched simulator (``mlp``, ``data``), and surrogate-assisted parameter fitting
(``fit``)."""
