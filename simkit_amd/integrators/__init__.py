from .backward_euler import backward_euler  # noqa: F401
