"""Minimal stand-in for the ``cprint`` package (a git+https dependency of the
reference that is not installable offline): ``c_print`` only prints."""


def c_print(*args, color=None, **kwargs):
    pass
