NAME          EDGES
ROWS
 N  obj
 E  r1
 L  r2
 G  r3
 E  r4
COLUMNS
    x1  obj  1   r1  1
    x1  r2   1
    x2  obj  -1  r1  1
    x2  r3   1
    x3  r2   2   r4  1
    x4  obj  0.5 r3  -1
    x5  r4   1
    x6  obj  2   r1  1
    x7  r2   1
    x8  obj  1   r3  1
RHS
    RHS1  obj  7  r1  4
    RHS1  r2   5  r3  1
RANGES
    RNG  r1  2   r2  3
    RNG  r3  -1.5  r4  -2
BOUNDS
 UP BND  x1  -1
 LO BND  x2  -3
 FX BND  x3  2.5
 FR BND  x4
 MI BND  x5
 UP BND  x6  4
 PL BND  x6
 BV BND  x7
QUADOBJ
    x1  x1  2
    x2  x1  1
    x2  x2  2
ENDATA
