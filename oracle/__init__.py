"""Oracle package (test infrastructure): CPU restatement of the reference path."""
