from .cuGillespie import CuGillespie
