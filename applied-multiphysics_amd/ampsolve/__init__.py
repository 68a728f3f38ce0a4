"""ampsolve: MI355X-native supernodal Cholesky / LDL^T factor and solves behind the reference's solver API."""
from .sparse import (  # noqa: F401
    DEFAULT_REAL_FACTOR,
    NumericFactor,
    SparseCholesky,
    SupernodalLDLT,
    Symbolic,
    amdOrder,
    ndOrder,
    analyze,
    checkSymbolic,
    factorNumeric,
    patternFingerprint,
    setDefaultRealFactor,
    solveNumeric,
    solveNumericMulti,
)
from .complex import ComplexSupernodalLDLT  # noqa: F401,E402
