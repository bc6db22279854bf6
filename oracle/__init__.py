"""ssq oracle: CPU restatement of the SpecSeq++ reduction path (checker only).

Test infrastructure. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
import this package, and only to check (or time) the product. The product package never
imports it.
"""
