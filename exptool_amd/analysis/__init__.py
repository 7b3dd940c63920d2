"""Bar-frame helpers used by the frozen-field path (exptool.analysis.pattern)."""
