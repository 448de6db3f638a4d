from .model import EMT6RoModel  # noqa: F401
