"""Adjoint-method gradients and batched variational circuits."""
