"""Minimal SBML reader with the slice of libsbml's API that SPN uses.

libsbml is not a dependency: this module parses SBML level 2/3 XML with
ElementTree and converts MathML kinetic laws into sympy expressions.  It
exposes the calls the reference front end makes
(mod/petri_net/SPN.py:50-150, examples/*_sim.py:14-21):

    reader = SBMLReader(); doc = reader.readSBML(path)
    doc.getNumErrors(); model = doc.getModel()
    model.getListOfSpecies() / getListOfReactions() / getListOfParameters()
    species.getId(), getInitialAmount(), getConstant(), getBoundaryCondition()
    reaction.getId(), getListOfReactants(), getListOfProducts(), getKineticLaw()
    ref.getSpecies(), ref.getStoichiometry(); kinetic_law.getMath()

``getMath()`` returns a :class:`MathExpr` holding a sympy expression over the
model's identifiers instead of a libsbml AST.
"""
import xml.etree.ElementTree as ET

import sympy

_MATHML = "{http://www.w3.org/1998/Math/MathML}"


def _local(tag):
    return tag.rsplit("}", 1)[-1]


def _children(elem, name):
    return [c for c in elem if _local(c.tag) == name]


def _child(elem, name):
    for c in elem:
        if _local(c.tag) == name:
            return c
    return None


class SBMLError(ValueError):
    pass


class MathExpr:
    """A kinetic law as a sympy expression over SBML identifiers."""

    def __init__(self, expr):
        self.expr = expr

    def names(self):
        return sorted(str(s) for s in self.expr.free_symbols)

    def __str__(self):
        return str(self.expr)


def _mathml_to_sympy(node):
    tag = _local(node.tag)
    if tag == "math":
        kids = list(node)
        if len(kids) != 1:
            raise SBMLError("<math> must hold exactly one expression")
        return _mathml_to_sympy(kids[0])
    if tag == "ci":
        return sympy.Symbol(node.text.strip())
    if tag == "cn":
        typ = node.get("type", "real")
        if typ == "e-notation":
            parts = [p.strip() for p in node.itertext() if p.strip()]
            return sympy.Float(parts[0]) * sympy.Integer(10) ** sympy.Integer(int(parts[1]))
        if typ == "rational":
            parts = [p.strip() for p in node.itertext() if p.strip()]
            return sympy.Rational(int(parts[0]), int(parts[1]))
        text = node.text.strip()
        if typ == "integer":
            return sympy.Integer(int(text))
        # an integral real literal stays exact (2 -> 2, 0.5 -> 1/2 is not assumed)
        try:
            return sympy.Integer(int(text))
        except ValueError:
            return sympy.Float(text)
    if tag == "apply":
        kids = list(node)
        op = _local(kids[0].tag)
        args = [_mathml_to_sympy(k) for k in kids[1:]]
        if op == "times":
            return sympy.Mul(*args)
        if op == "plus":
            return sympy.Add(*args)
        if op == "minus":
            if len(args) == 1:
                return -args[0]
            if len(args) == 2:
                return args[0] - args[1]
            raise SBMLError("<minus/> takes one or two arguments")
        if op == "divide":
            if len(args) != 2:
                raise SBMLError("<divide/> takes two arguments")
            return args[0] / args[1]
        if op == "power":
            if len(args) != 2:
                raise SBMLError("<power/> takes two arguments")
            return args[0] ** args[1]
        if op == "exp":
            return sympy.exp(args[0])
        if op == "ln":
            return sympy.log(args[0])
        raise SBMLError("MathML operator <%s/> is not supported" % op)
    raise SBMLError("MathML element <%s> is not supported" % tag)


class Species:
    def __init__(self, e):
        self._id = e.get("id")
        self._name = e.get("name", self._id)
        amt = e.get("initialAmount")
        if amt is None:
            amt = e.get("initialConcentration", "0")
        self._amount = float(amt)
        self._constant = e.get("constant", "false").lower() == "true"
        self._boundary = e.get("boundaryCondition", "false").lower() == "true"

    def getId(self):
        return self._id

    def getName(self):
        return self._name

    def getInitialAmount(self):
        return self._amount

    def getConstant(self):
        return self._constant

    def getBoundaryCondition(self):
        return self._boundary


class Parameter:
    def __init__(self, e):
        self._id = e.get("id")
        self._value = float(e.get("value", "0"))

    def getId(self):
        return self._id

    def getValue(self):
        return self._value


class SpeciesReference:
    def __init__(self, e):
        self._species = e.get("species")
        self._stoich = float(e.get("stoichiometry", "1"))

    def getSpecies(self):
        return self._species

    def getStoichiometry(self):
        return self._stoich


class KineticLaw:
    def __init__(self, e):
        m = _child(e, "math")
        if m is None:
            raise SBMLError("kineticLaw without <math>")
        self._math = MathExpr(_mathml_to_sympy(m))
        self._locals = {}
        for lst in ("listOfParameters", "listOfLocalParameters"):
            node = _child(e, lst)
            if node is not None:
                for p in node:
                    self._locals[p.get("id")] = float(p.get("value", "0"))

    def getMath(self):
        return self._math

    def getLocalParameters(self):
        return dict(self._locals)


class Reaction:
    def __init__(self, e):
        self._id = e.get("id")
        self._name = e.get("name", self._id)
        r = _child(e, "listOfReactants")
        p = _child(e, "listOfProducts")
        self._reactants = [SpeciesReference(x) for x in (r if r is not None else [])
                           if _local(x.tag) == "speciesReference"]
        self._products = [SpeciesReference(x) for x in (p if p is not None else [])
                          if _local(x.tag) == "speciesReference"]
        kl = _child(e, "kineticLaw")
        self._kl = KineticLaw(kl) if kl is not None else None

    def getId(self):
        return self._id

    def getName(self):
        return self._name

    def getListOfReactants(self):
        return list(self._reactants)

    def getListOfProducts(self):
        return list(self._products)

    def getKineticLaw(self):
        if self._kl is None:
            raise SBMLError("reaction %s has no kineticLaw" % self._id)
        return self._kl


class Model:
    def __init__(self, e):
        self._id = e.get("id")
        self._name = e.get("name", self._id)
        self._compartments = {}
        node = _child(e, "listOfCompartments")
        for c in (node if node is not None else []):
            self._compartments[c.get("id")] = float(c.get("size", c.get("volume", "1")))
        node = _child(e, "listOfSpecies")
        self._species = [Species(s) for s in (node if node is not None else [])
                         if _local(s.tag) == "species"]
        node = _child(e, "listOfParameters")
        self._params = [Parameter(p) for p in (node if node is not None else [])
                        if _local(p.tag) == "parameter"]
        node = _child(e, "listOfReactions")
        self._reactions = [Reaction(r) for r in (node if node is not None else [])
                           if _local(r.tag) == "reaction"]

    def getId(self):
        return self._id

    def getName(self):
        return self._name

    def getListOfSpecies(self):
        return list(self._species)

    def getListOfParameters(self):
        return list(self._params)

    def getListOfReactions(self):
        return list(self._reactions)

    def getCompartmentSizes(self):
        return dict(self._compartments)


class Document:
    def __init__(self, model, errors):
        self._model = model
        self._errors = errors

    def getNumErrors(self):
        return len(self._errors)

    def getErrors(self):
        return list(self._errors)

    def getModel(self):
        return self._model


class SBMLReader:
    def readSBML(self, path):
        with open(path, "rb") as f:
            return self.readSBMLFromString(f.read())

    def readSBMLFromString(self, text):
        try:
            root = ET.fromstring(text)
        except ET.ParseError as exc:
            return Document(None, [str(exc)])
        if _local(root.tag) != "sbml":
            return Document(None, ["root element is not <sbml>"])
        m = _child(root, "model")
        if m is None:
            return Document(None, ["no <model> element"])
        try:
            return Document(Model(m), [])
        except SBMLError as exc:
            return Document(None, [str(exc)])


def read_model(path):
    """Read an SBML file and return its model, raising on errors."""
    doc = SBMLReader().readSBML(path)
    if doc.getNumErrors() > 0:
        raise SBMLError("%d errors in SBML file %s: %s" % (doc.getNumErrors(), path, doc.getErrors()))
    return doc.getModel()
