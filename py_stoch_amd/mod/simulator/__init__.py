from .simulator import Simulator, StochasticSimulator
