"""Namespace of the MI355X-native regulariser package (mirrors the reference's ``ccpi``)."""
