"""Trajectory saving in the reference's npz format (ds_saving/saving.py)."""
