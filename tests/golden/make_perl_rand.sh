#!/bin/sh
# Writes Perl's own rand() stream after srand(12345) -- the generator
# nbkc_normalize.pl (srand at :69, rand(1) at :111) draws from.  Runs the
# system perl interpreter only; no reference script is executed.
perl -e 'srand(12345); printf("%.17g\n", rand(1)) for 1..2000' > perl_rand_12345.txt
perl -e 'printf("%s\n", $^V)' > perl_version.txt
